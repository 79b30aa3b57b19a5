/*
 * lsoracle.h -- CPU restatement of the reference's sparse level-set segmentation
 * path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the correctness oracle for the HIP engine.  Only the test
 * suite, the smoke check in __graft_entry__.py and the cpu_baseline leg of
 * bench.py may load it, and only as the checker or the timed CPU baseline.
 * The product path (the HIP library and its C++/Python hosts) never calls it.
 *
 * Every function restates one piece of the reference (file:line cited at the
 * definition in lsoracle.c).  Reference paths are relative to
 * code/gpuseg/code/ of sjpatel21/AWorkEfficientGpuAlgorithmForLevelSetSegmentation.
 *
 * Data layout (shared with the HIP engine so states compare byte for byte):
 *   - padded dims X, Y, Z are multiples of 16 (Segmenter.cpp:1136-1143);
 *   - every per-voxel array (phi, source, export byte, scratch) is indexed by
 *     the 16^3-tiled index tau(x,y,z) of Coordinates.cu:65-96;
 *   - an active list holds tau values (the reference packs 10/10/10-bit
 *     coordinates, Coordinates.cu:186-208; tau is a bijection, so list order
 *     and membership are preserved).
 *
 * Semantics implemented ("REF" mode, the parity target): the shipped update
 * rule (UpdateLevelSetVolume.cu:222-342), the shipped marking predicate
 * (UpdateActiveElements.cu:75-103) including the skip of list positions
 * p % 4 == 3 (UpdateActiveElementsVolume.cu:82,329), first-wins duplicate
 * removal over the seven candidate lists (FilterDuplicates.cu:16-193), stable
 * compaction (Cuda.cpp:699-758), the few/many path switch
 * (Segmenter.cpp:1363-1369) and ping-pong phi buffers with no copy-through
 * (Segmenter.cpp:841,1339-1351).  Scratch state is clean every iteration (the
 * stale-flag artefacts of the shipped driver are not replayed).  "CLEAN" mode
 * drops the p % 4 == 3 skip (ablation only).
 */
#ifndef LSORACLE_H
#define LSORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* source element types (VolumeDesc numBytesPerVoxel x isSigned, plus f32) */
enum { LSO_U8 = 0, LSO_I8 = 1, LSO_U16 = 2, LSO_I16 = 3, LSO_U32 = 4, LSO_I32 = 5, LSO_F32 = 6 };

/* candidate list order of UpdateActiveElementsVolume.cu:168-174 */
enum { LSO_NZ = 0, LSO_NY = 1, LSO_NX = 2, LSO_SELF = 3, LSO_PX = 4, LSO_PY = 5, LSO_PZ = 6 };

enum { LSO_MODE_REF = 0, LSO_MODE_CLEAN = 1 };

/* ---- index helpers -------------------------------------------------------- */
uint32_t lso_tau(int32_t X, int32_t Y, int32_t x, int32_t y, int32_t z);
void     lso_untau(int32_t X, int32_t Y, uint32_t t, int32_t* x, int32_t* y, int32_t* z);
/* neighbour in list direction dir; out-of-bounds returns t itself */
uint32_t lso_neighbor(int32_t X, int32_t Y, int32_t Z, uint32_t t, int32_t dir);
int32_t  lso_source_bytes(int32_t dtype);
uint64_t lso_align256(uint64_t n);
int32_t  lso_path_is_few(uint64_t n, uint64_t N);

/* ---- stateless kernels (arrays are tiled, length N = X*Y*Z) ---------------- */
void lso_update_f32(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src,
                    const float* R, float* W, int8_t* E, uint8_t* flags,
                    const uint32_t* V, int64_t n,
                    float T, float eps, float alpha, float dt, int32_t threads);
void lso_update_f64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src,
                    const double* R, double* W, int8_t* E, uint8_t* flags,
                    const uint32_t* V, int64_t n,
                    double T, double eps, double alpha, double dt, int32_t threads);
void lso_mark(int32_t X, int32_t Y, int32_t Z, const float* R, const uint8_t* flags,
              const uint32_t* V, int64_t n, int32_t mode, uint8_t* masks, int32_t threads);
int64_t lso_compact_few(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                        const uint8_t* masks, uint32_t* out);
int64_t lso_compact_many(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                         const uint8_t* masks, uint32_t* out);
/* the same compactions in parallel with identical output; tag / flag: N bytes
   of zeroed scratch (zero again on return), keep: n bytes of scratch */
int64_t lso_compact_few_par(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                            const uint8_t* masks, uint32_t* out, uint8_t* tag, uint8_t* keep,
                            int32_t threads);
int64_t lso_compact_many_par(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                             const uint8_t* masks, uint32_t* out, uint8_t* flag, int32_t threads);

/* ---- stateful engine ------------------------------------------------------ */
typedef struct lso_state lso_state;

lso_state* lso_create(int32_t nx, int32_t ny, int32_t nz, int32_t dtype, const void* src_linear,
                      float z_aniso);
void       lso_destroy(lso_state* s);
void       lso_dims(const lso_state* s, int32_t* X, int32_t* Y, int32_t* Z);
/* phi0 over the unpadded box (x fastest); padded voxels get 1.0; V0 = {|phi0| < band} */
/* fixed_point != 0: phi stored as the reference's LEVEL_SET_CHAR (call before
   initialising phi) */
void       lso_set_phi_storage(lso_state* s, int32_t fixed_point);
int32_t    lso_set_phi(lso_state* s, const float* phi0_linear, float band);
/* reference seed initialisation: two init passes + gradient keep-flags */
int32_t    lso_seed_sphere(lso_state* s, float cx, float cy, float cz, uint32_t radius);
/* run iters iterations; per-iteration n (active), c (candidates), m (new active) */
int32_t    lso_iterate(lso_state* s, float T, float eps, float alpha, float dt, int32_t mode,
                       int32_t iters, int32_t threads,
                       uint32_t* n_hist, uint64_t* c_hist, uint32_t* m_hist);
int64_t    lso_active_count(const lso_state* s);
int64_t    lso_get_active(const lso_state* s, uint32_t* out, int64_t cap);
/* raw tiled state: read buffer (current), write buffer, export byte, list */
void       lso_get_state(const lso_state* s, float* phi_read, float* phi_write, int8_t* E);
int32_t    lso_set_state(lso_state* s, const float* phi_read, const float* phi_write,
                         const int8_t* E, const uint32_t* V, int64_t n);
/* unpadded, linear (x fastest) outputs */
void       lso_get_phi(const lso_state* s, int32_t which, float* out_linear);
void       lso_get_mask(const lso_state* s, uint8_t* out_linear);
void       lso_reinit_active(lso_state* s);
void       lso_iterate_reinit(lso_state* s, float T, float eps, float alpha, float dt);
void       lso_freeze(lso_state* s);
int64_t    lso_auto_parameters(const lso_state* s, float* T, float* eps);
void       lso_clear(lso_state* s, int32_t all);
/* LITERAL mode (diagnostic): replay the shipped driver's scratch volumes with
   their stale contents (see lsoracle.c); begin from the current state, then
   iterate; events: 3 per iteration (stale valid flags compacted, candidates
   meeting a stale tag, repeated targets within a list) */
int32_t    lso_literal_begin(lso_state* s, int32_t initial_flags);
int32_t    lso_iterate_literal(lso_state* s, float T, float eps, float alpha, float dt, int32_t iters,
                               int32_t threads, uint32_t* n_hist, uint64_t* c_hist, uint32_t* m_hist,
                               uint64_t* events);
/* union of two export volumes (InitializeLevelSetVolume.cu:151-169) */
void       lso_add_export(int8_t* result, const int8_t* add, int64_t N);

/* ---- differentiable form (lsoracle_adjoint.c) ----------------------------- */
void lso_decisions_f32(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, const float* R,
                       float T, float eps, float alpha, float dt, const uint32_t* V, int64_t n, uint8_t* dec);
void lso_replay_f64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, double T, double eps,
                    double alpha, double dt, int32_t iters, const uint32_t* n, const uint32_t* lists,
                    const uint8_t* dec, const double* phi0, double* phi_out, double* A_hist);
void lso_adjoint_f64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, double T, double eps,
                     double alpha, double dt, int32_t iters, const uint32_t* n, const uint32_t* lists,
                     const uint8_t* dec, const double* A_hist, const double* dphiK, double* dphi0, double* dtheta);

#ifdef __cplusplus
}
#endif

#endif
