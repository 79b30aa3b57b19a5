/*
 * lsseg.h -- C ABI of the MI355X (gfx950) sparse level-set segmentation engine.
 *
 * This is the drop-in boundary for the reference's GPU segmentation path.  The
 * reference exposes the same path at two levels, and this ABI replaces both:
 *
 *   - the C++ entry point `Segmenter` (code/gpuseg/code/Segmenter.hpp:40-73),
 *     which owns all device buffers and drives the per-iteration loop
 *     (Segmenter.cpp:759-1051);
 *   - the `extern "C"` kernel launchers behind it
 *     (code/gpuseg/code/cuda/CudaKernels.hpp:9-95), which take raw device
 *     pointers plus implicitly bound texture references.
 *
 * One opaque context owns every device buffer (both phi buffers, the export
 * byte, the active lists, scratch and counters), so there is no global state and
 * no texture binding.  Every call returns a status and never aborts; the text of
 * the last error is available from lsseg_last_error().  A context is externally
 * synchronised; separate contexts may be used from separate threads.
 *
 * Conventions
 *   - Volumes crossing the ABI are linear, x fastest, unpadded (nx*ny*nz).
 *     Internally each axis is padded to a multiple of 16
 *     (Segmenter.cpp:1131-1204) and split into the reference's 16^3 tiles
 *     (Coordinates.cu:65-96).
 *   - Active-list entries and raw padded state cross the ABI in the reference's
 *     tiled index tau of the padded volume:
 *     tau = ((tz*TY + ty)*TX + tx)*4096 + (lz*16 + ly)*16 + lx
 *     with (tx, lx) = (x/16, x%16) and likewise for y, z; TX = X/16, TY = Y/16.
 *     Device memory holds the voxels of a tile in Morton order instead
 *     (tile*4096 + m, bit 3k of m = bit k of lx, 3k+1 of ly, 3k+2 of lz); only
 *     lsseg_get_device_buffers exposes that layout.
 *   - The caller owns host arrays; the context copies them in and out.
 */
#ifndef LSSEG_LSSEG_H
#define LSSEG_LSSEG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsseg_ctx lsseg_ctx;
typedef struct lsseg_batch lsseg_batch;
typedef struct lsseg_record lsseg_record;

typedef enum lsseg_status {
    LSSEG_OK = 0,
    LSSEG_ERR_INVALID_ARG = 1,   /* null pointer, bad size, bad parameter */
    LSSEG_ERR_OUT_OF_MEMORY = 2, /* device or host allocation failed */
    LSSEG_ERR_DEVICE = 3,        /* HIP runtime error or device-side error flag */
    LSSEG_ERR_CAPACITY = 4,      /* caller buffer too small */
    LSSEG_ERR_STATE = 5,         /* call not valid in the current state */
    LSSEG_ERR_UNSUPPORTED = 6    /* volume too large, unsupported type */
} lsseg_status;

/* Source element type: VolumeDesc::numBytesPerVoxel x isSigned
   (VolumeDesc.hpp:12-38, Textures.cu:5-10), plus 32-bit float sources. */
typedef enum lsseg_dtype {
    LSSEG_U8 = 0, LSSEG_I8 = 1, LSSEG_U16 = 2, LSSEG_I16 = 3,
    LSSEG_U32 = 4, LSSEG_I32 = 5, LSSEG_F32 = 6
} lsseg_dtype;

/* Storage of phi.  LSSEG_PHI_F32 keeps FP32 values (the parity target with the
   FP64 oracle).  LSSEG_PHI_I8 is the reference's shipped LEVEL_SET_CHAR
   representation (Config.hpp:72-80): stored as floor(127 phi)
   (UpdateLevelSetVolume.cu:316-318), read as s / 127 with -128 -> -1 (the
   normalised-float texture read, Textures.cu:12); 1 byte instead of 4 per
   value.  The differentiable form needs LSSEG_PHI_F32. */
typedef enum lsseg_phi_storage { LSSEG_PHI_F32 = 0, LSSEG_PHI_I8 = 1 } lsseg_phi_storage;

/* Replaces VolumeDesc (VolumeDesc.hpp:12-38): dims, type, z anisotropy, plus
   the phi storage (zero-initialised: FP32). */
typedef struct lsseg_volume_desc {
    uint32_t nx, ny, nz;
    int32_t dtype;       /* lsseg_dtype */
    float z_aniso;       /* VolumeDesc::zAnisotropy, used by the seed sphere */
    int32_t phi_storage; /* lsseg_phi_storage */
} lsseg_volume_desc;

typedef enum lsseg_mode {
    LSSEG_MODE_REF = 0,   /* shipped rule: list positions p % 4 == 3 never emit
                             (UpdateActiveElementsVolume.cu:82,329) */
    LSSEG_MODE_CLEAN = 1  /* every position emits (ablation) */
} lsseg_mode;

/* Replaces the Segmenter's global parameters growLevel, growWindow,
   curvatureInfluence and timeStep (Segmenter.cpp:34-40, passed at :811-822).
   For integer sources T and eps are truncated to int as the reference does
   (Segmenter.cpp:817-818); for float sources they are used as given. */
typedef struct lsseg_params {
    float T;       /* target intensity (growLevel) */
    float eps;     /* window (growWindow, maxDistanceBeforeShrink) */
    float alpha;   /* curvature weight (curvatureInfluence) */
    float dt;      /* time step (timeStep); dt < 0 flips the data term */
    int32_t mode;  /* lsseg_mode */
} lsseg_params;

/* Per-call statistics (sums over the iterations of one lsseg_iterate call). */
typedef struct lsseg_stats {
    uint64_t iterations;       /* iterations run */
    uint64_t sum_active;       /* sum of n_k: active-voxel updates */
    uint64_t sum_candidates;   /* sum of c_k: candidate emissions before dedup */
    uint64_t sum_new_active;   /* sum of m_k: entries written to the next lists */
    uint64_t active_final;     /* n after the last iteration */
    uint64_t few_path_iters;   /* iterations that took the sparse ("few") path */
    uint64_t many_path_iters;  /* iterations that took the dense ("many") path */
    double   device_ms;        /* device time of the call (HIP events) */
    uint64_t empty_skipped;    /* trailing iterations over an empty list that were not
                                  launched: they change nothing but the counters, which
                                  are set as if they had run */
} lsseg_stats;

/* ---- lifetime ----------------------------------------------------------- */

/* Replaces Segmenter::LoadVolume (Segmenter.cpp:110-285): pads, tiles and
   uploads the source; allocates every buffer; phi := 1, export := 127. */
lsseg_status lsseg_create(const lsseg_volume_desc* desc, const void* host_volume,
                          lsseg_ctx** out_ctx);

/* Replaces Segmenter::UnloadVolume / ~Segmenter (Segmenter.cpp:97-108, 287-340). */
lsseg_status lsseg_destroy(lsseg_ctx* ctx);

/* Text of the last error on this context ("" if none).  With ctx == NULL,
   the last error of a failed lsseg_create on this thread. */
const char* lsseg_last_error(const lsseg_ctx* ctx);

/* Padded dims (multiples of 16) of the context. */
lsseg_status lsseg_get_dims(const lsseg_ctx* ctx, uint32_t* X, uint32_t* Y, uint32_t* Z);

/* ---- initialisation ----------------------------------------------------- */

/* Replaces Segmenter::InitializeSeed (Segmenter.cpp:491-595) with the reference
   semantics: the centre is truncated to unsigned ints, two init passes with a
   buffer swap after each union the sphere ramp into phi
   (InitializeLevelSetVolume.cu:55-98), then the active list is every voxel
   whose 6-neighbour differences exceed 0.001
   (InitializeActiveElementsVolume.cu:37-74) in ascending tiled order. */
lsseg_status lsseg_seed_sphere(lsseg_ctx* ctx, float cx, float cy, float cz, uint32_t radius);

/* Injects phi0 (unpadded, linear; host or device memory) into both phi buffers, export := floor(127
   phi0) saturated, padded voxels := 1.0; active list := {|phi0| < band} in
   ascending tiled order.  No reference counterpart: the benchmark
   configurations start from a signed-distance phi0. */
lsseg_status lsseg_set_phi(lsseg_ctx* ctx, const float* phi0, float band);

/* ---- iteration ---------------------------------------------------------- */

/* Replaces the sub-step loop of Segmenter::RequestUpdateSegmentation
   (Segmenter.cpp:772-1036) and the kernel launchers it calls
   (CudaUpdateLevelSetVolumeAsync, CudaOutputNewActiveElements,
   CudaUpdateActiveElementsVolume, CudaMemSetCharSparse, CudaFilterDuplicates,
   CudppCompact, CudaMemSetIntSparse; CudaKernels.hpp:25-87, Cuda.cpp:699-758).
   Runs `iters` iterations on the context's stream with no host round trip
   between them and synchronises once at return.  stats may be NULL. */
lsseg_status lsseg_iterate(lsseg_ctx* ctx, const lsseg_params* params, int32_t iters,
                           lsseg_stats* stats);

/* One iteration whose marking is replaced by a restart of the active set: the
   update runs with the given parameters, then the active list becomes every
   voxel whose 6-neighbour differences on the new phi exceed 0.001, in ascending
   tiled order.  Replaces the parameter-change branch of
   Segmenter::RequestUpdateSegmentation (Segmenter.cpp:847-876, taken when
   ShouldInitializeActiveElements, :1371-1374, holds). */
lsseg_status lsseg_iterate_reinit(lsseg_ctx* ctx, const lsseg_params* params);

/* Per-iteration counters of the last lsseg_iterate call: n_k (active), c_k
   (candidates), m_k (new active).  Any pointer may be NULL; with all three NULL
   the call only reports *count.  Returns LSSEG_ERR_CAPACITY (after copying cap
   entries) if cap < *count. */
lsseg_status lsseg_get_iteration_counts(const lsseg_ctx* ctx, uint32_t* n, uint64_t* c,
                                        uint32_t* m, uint64_t cap, uint64_t* count);

/* ---- readback ----------------------------------------------------------- */

/* Replaces Segmenter::GetActiveElementCount (Segmenter.hpp:66). */
lsseg_status lsseg_get_active_count(const lsseg_ctx* ctx, uint64_t* n);

/* Active list in list order (tiled indices); *n receives the full count.
   Returns LSSEG_ERR_CAPACITY (after copying cap entries) if cap < n. */
lsseg_status lsseg_get_active(const lsseg_ctx* ctx, uint32_t* out, uint64_t cap, uint64_t* n);

/* phi over the unpadded box into out (host or device memory); which = 0: the
   read buffer (latest values at active voxels), which = 1: the other buffer. */
lsseg_status lsseg_get_phi(const lsseg_ctx* ctx, int32_t which, float* out);

/* Replaces Segmenter::GetSaveableSegmentation (Segmenter.cpp:1216-1286) plus
   UnpadVolume (:1288-1337): out[i] = 1 where the export byte is <= 0, else 0,
   over the unpadded box.  With the frozen segmentation included
   (lsseg_freeze), the export is first unioned with it as AddLevelSetKernel does
   (InitializeLevelSetVolume.cu:151-169). */
lsseg_status lsseg_get_mask(const lsseg_ctx* ctx, uint8_t* out);

/* Replaces Segmenter::FreezeCurrentSegmentation (Segmenter.cpp:659-679):
   frozen := union(frozen, export). */
lsseg_status lsseg_freeze(lsseg_ctx* ctx);

/* Replaces Segmenter::CalculateSegmentationParameters (Segmenter.cpp:357-478):
   over the voxels whose export byte is <= 0 (the current segmentation, without
   the frozen one), *T := mean and *eps := sample standard deviation of the
   source values; *count (may be NULL) := the number of such voxels.  The sums
   run in double; the reference sums sequentially in float, so results differ
   by that rounding only.  Fails with LSSEG_ERR_STATE below two voxels. */
lsseg_status lsseg_auto_parameters(lsseg_ctx* ctx, float* T, float* eps, uint64_t* count);

/* Replaces Segmenter::ClearCurrentSegmentation (Segmenter.cpp:597-640): phi := 1,
   export := 127, empty active list.  all != 0 also clears the frozen
   segmentation (ClearAllSegmentations, :641-657). */
lsseg_status lsseg_clear(lsseg_ctx* ctx, int32_t all);

/* ---- raw state (checkpoint / resume, parity tests) ---------------------- */

/* Raw padded tiled state: phi read buffer, phi write buffer, export byte
   (N = X*Y*Z entries each; any pointer may be NULL) and the active list.  With
   LSSEG_PHI_I8 storage phi travels as the floats s / 127 it reads as, and
   lsseg_set_state stores round(127 v). */
lsseg_status lsseg_get_state(const lsseg_ctx* ctx, float* phi_read, float* phi_write,
                             int8_t* export_bytes, uint32_t* list, uint64_t list_cap,
                             uint64_t* n);
lsseg_status lsseg_set_state(lsseg_ctx* ctx, const float* phi_read, const float* phi_write,
                             const int8_t* export_bytes, const uint32_t* list, uint64_t n);

/* Device-resident checkpoint of the complete segmentation state (both phi
   buffers, export byte, active list) in slot 0..LSSEG_MAX_CHECKPOINTS-1; the slot
   is allocated on first use.  Restore makes the saved state current again.  No
   reference counterpart (the reference keeps no engine checkpoint); used for
   resumable runs, repeated timed runs from one resident input, and the start
   state that lsseg_fwd_record records again from when its first size guess
   was short. */
#define LSSEG_MAX_CHECKPOINTS 64
lsseg_status lsseg_checkpoint_save(lsseg_ctx* ctx, int32_t slot);
lsseg_status lsseg_checkpoint_restore(lsseg_ctx* ctx, int32_t slot);

/* Device time of the iteration kernels.  Every launch carries its own start
   and stop HIP events on the context's stream (hipExtLaunchKernel), which take
   the kernel's start and end timestamps.  An lsseg_iterate call of `iters`
   (>= 1) iterations makes 3 * iters + 1 launches: the update, marking and keep
   step (duplicate removal) of each iteration in turn, where the update of
   every iteration after the first also runs the previous iteration's
   compaction, then the last iteration's compaction.  Under the environment
   setting LSSEG_FUSE=0 every iteration is four launches (update, marking,
   keep, compaction) instead, with identical results. */
typedef struct lsseg_kernel_times {
    double total_ms;     /* sum over all iteration launches */
    double first_ms;     /* the first launch */
    double last_ms;      /* the last launch (the final compaction) */
    uint64_t launches;   /* 3 * L + 1 for L launched iterations */
} lsseg_kernel_times;

/* lsseg_iterate with per-launch event timing (launches issued one by one, no
   graph).  Results and statistics are identical to lsseg_iterate.  launch_ms
   (may be NULL) receives up to launch_cap per-launch times in launch order.
   Like lsseg_iterate it stops launching once the list is empty (checked every
   64 iterations), so L <= iters iterations are launched; stats->empty_skipped
   counts the rest. */
lsseg_status lsseg_iterate_profiled(lsseg_ctx* ctx, const lsseg_params* params, int32_t iters,
                                    lsseg_stats* stats, lsseg_kernel_times* times,
                                    float* launch_ms, uint64_t launch_cap);

/* Device pointers of the phi buffers (read, write; float or int8_t by the
   context's phi storage; padded, tiles of 16^3 voxels in Morton order, see
   the conventions above) and the stream, for callers that keep data resident
   in HBM (e.g. a torch wrapper). */
lsseg_status lsseg_get_device_buffers(const lsseg_ctx* ctx, void** phi_read, void** phi_write,
                                      void** stream);

/* ---- one run that allocates nothing -------------------------------------

   No counterpart in the reference.  For callers that cannot allocate device
   memory: code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc.  One call does what lsseg_create + lsseg_set_phi +
   lsseg_iterate + lsseg_get_phi(ctx, 0, .) + lsseg_get_mask do, with the same
   kernels and bit-identical results, but calls no hipMalloc, hipFree or other
   allocating routine: every buffer except the caller's inputs and outputs
   lives in one static device array of LSSEG_NOALLOC_SCRATCH_BYTES bytes per
   GPU, which the library holds until the process ends.

   src (unpadded source, x fastest, desc->dtype elements), phi0 (unpadded
   float) are device inputs; phi_out (unpadded float, the final read buffer)
   and mask_out (unpadded bytes, export <= 0) are device outputs; all four in
   device memory of the current device.  The work is queued on `stream` (a
   hipStream_t; NULL: the null stream).  With stats == NULL the call only
   queues kernels and returns (no wait, no memset, no host copy: it can be
   captured in a graph); with stats it waits for the run, reports a device
   error flag, and fills stats (device_ms and empty_skipped stay 0: every
   iteration is launched).
   Limits: at most LSSEG_NOALLOC_MAX_VOXELS padded voxels (each axis padded to
   a multiple of 16) and LSSEG_NOALLOC_MAX_ITERS iterations; larger requests
   fail with LSSEG_ERR_UNSUPPORTED.  Every call on a GPU uses that GPU's one
   array and leaves its contents to the next call: calls that overlap on one
   GPU (other streams, other threads) race and must be serialised by the
   caller.  Errors: lsseg_last_error(NULL). */
#define LSSEG_NOALLOC_MAX_VOXELS (1ull << 21)  /* 128^3 */
#define LSSEG_NOALLOC_MAX_ITERS 16384ull
#define LSSEG_NOALLOC_SCRATCH_BYTES (47ull * LSSEG_NOALLOC_MAX_VOXELS + (2ull << 20))
lsseg_status lsseg_segment_noalloc(const lsseg_volume_desc* desc, const void* src, const float* phi0, float band,
                                   const lsseg_params* params, int32_t iters, float* phi_out, uint8_t* mask_out,
                                   void* stream, lsseg_stats* stats);

/* ---- batched form --------------------------------------------------------

   No reference counterpart: the reference segments one volume at a time.  A
   batch advances many independent volumes of uneven size, each with its own
   parameters, with one launch per step (update, marking, compaction) for the
   whole batch per iteration; runs of iterations replay one captured hipGraph.
   Each volume's results are bit-identical to a solo lsseg_iterate run of the
   same volume.  All volumes share the batch's stream and one source type. */

#define LSSEG_MAX_BATCH 1024

/* Creates `count` volumes (descs[i], host_volumes[i]); all descs must share
   one dtype.  On failure *out is NULL and lsseg_batch_last_error(NULL) says why. */
lsseg_status lsseg_batch_create(const lsseg_volume_desc* descs, const void* const* host_volumes,
                                uint32_t count, lsseg_batch** out);
lsseg_status lsseg_batch_destroy(lsseg_batch* batch);
const char* lsseg_batch_last_error(const lsseg_batch* batch);
lsseg_status lsseg_batch_count(const lsseg_batch* batch, uint32_t* count);

/* The context of volume `index`, for initialisation and readback with the
   single-volume calls (lsseg_set_phi, lsseg_seed_sphere, lsseg_get_mask,
   lsseg_get_active, lsseg_get_phi, lsseg_freeze, lsseg_clear, checkpoints ...).
   It is owned by the batch: lsseg_destroy on it fails with LSSEG_ERR_STATE. */
lsseg_status lsseg_batch_volume(lsseg_batch* batch, uint32_t index, lsseg_ctx** out);

/* `iters` iterations of every volume; params[i] applies to volume i; stats
   (may be NULL) receives one lsseg_stats per volume (device_ms is the batch's). */
lsseg_status lsseg_batch_iterate(lsseg_batch* batch, const lsseg_params* params, int32_t iters,
                                 lsseg_stats* stats);

/* Checkpoint every volume into / from its slot `slot` (lsseg_checkpoint_*). */
lsseg_status lsseg_batch_checkpoint_save(lsseg_batch* batch, int32_t slot);
lsseg_status lsseg_batch_checkpoint_restore(lsseg_batch* batch, int32_t slot);

/* Launch shape and widths of the batched kernels: workgroups of the update
   and marking launches and of the keep and compaction launches (more than the
   volume count, at most the resident count), traversal positions per thread
   of the keep step (1 or 2; 0 in set_tuning: unchanged) and list positions per
   few-path compaction tile (256 or 1024; 0: unchanged). */
typedef struct lsseg_tuning {
    uint32_t update_blocks;
    uint32_t compact_blocks;
    double ms;             /* autotune: device time of the chosen tuning */
    uint32_t candidates;   /* autotune: tunings timed (0 when taken from the cache) */
    uint32_t keep_items;
    uint32_t compact_tile;
    uint32_t cached;       /* autotune: 1 when the result came from the tuning cache */
} lsseg_tuning;

lsseg_status lsseg_batch_get_tuning(const lsseg_batch* batch, lsseg_tuning* out);
lsseg_status lsseg_batch_set_tuning(lsseg_batch* batch, const lsseg_tuning* tuning);

/* Times `iters` iterations of the batch from its current state for candidate
   tunings and keeps the fastest, by coordinate descent: 12 launch shapes at
   the default widths, then the keep step's positions per thread, then the
   compaction tile; restores the state (uses checkpoint slot
   LSSEG_MAX_CHECKPOINTS - 1 of every volume).  Results are cached per (GPU
   architecture, source type, phi storage, volume count, total voxels to the
   next power of two, iters): in the process, and in the file LSSEG_TUNING_CACHE
   names if it is set; a cached result is applied without timing
   (LSSEG_AUTOTUNE_FORCE=1 times again). */
lsseg_status lsseg_batch_autotune(lsseg_batch* batch, const lsseg_params* params, int32_t iters,
                                  lsseg_tuning* out);

/* The same for one context (the reference hard-codes its launch shapes,
   UpdateLevelSetVolume.cu:42-55, Cuda.cpp:699-758).  Fields: update_blocks =
   workgroups of the update and marking launches, compact_blocks = workgroups of
   the keep, compaction and fused launches (multiples of 8, at least 8 and 16),
   keep_items = traversal positions per thread of the keep step (1 or 2),
   compact_tile = 256: few-path compactions of lists up to the context's
   small-list bound (LSSEG_SMALL_TILE_N at creation) take tiles of 256
   positions, 1024: every compaction takes tiles of 1024; 0 in a field of
   set_tuning leaves it unchanged.  Results are identical under every tuning.
   lsseg_autotune times `iters` iterations from the current state for 15 launch
   shapes, then the other keep width, then the other compaction tile, keeps
   the fastest and restores the state (checkpoint slot LSSEG_MAX_CHECKPOINTS -
   1); it caches like lsseg_batch_autotune, per (GPU architecture, source
   type, phi storage, voxels to the next power of two, iters). */
lsseg_status lsseg_get_tuning(const lsseg_ctx* ctx, lsseg_tuning* out);
lsseg_status lsseg_set_tuning(lsseg_ctx* ctx, const lsseg_tuning* tuning);
lsseg_status lsseg_autotune(lsseg_ctx* ctx, const lsseg_params* params, int32_t iters, lsseg_tuning* out);

/* ---- differentiable form -------------------------------------------------

   No reference counterpart.  The adjoint of the update lets a caller tune T,
   eps, alpha and phi0 against a loss on the final phi.  The active set and the
   branch decisions of every update (the two clamps, the connectivity gate, a
   vanishing gradient, the side of |I - T|) are recorded per iteration and held
   fixed: the map from
   (phi0, T, eps, alpha) to the final phi is then smooth, lsseg_fwd_replay
   evaluates it for any inputs, and lsseg_bwd is its exact adjoint (up to FP32
   rounding).  The record also keeps, per update, the value of the write buffer
   it overwrote, and the final buffer pair, so lsseg_bwd steps the buffers back
   from the end without checkpoints or recomputation (9 bytes per recorded
   entry and 8 per voxel in all).  For integer sources T and eps are truncated
   (Segmenter.cpp:817-818), so their derivatives are zero.

   Pointers marked "host or device" may point to either (unified addressing);
   volumes are linear, x fastest, unpadded. */

/* Runs `iters` iterations of lsseg_iterate from the current state (which
   lsseg_set_phi normally made: both buffers = phi0) and records them.  The
   context ends in the same state as after lsseg_iterate.  One run records; the
   record's size is known only on the device, so the run starts from a guess
   (the context's last record, scaled) and runs again with the exact size when
   the guess was short.  Uses checkpoint slot LSSEG_MAX_CHECKPOINTS - 1.
   Recorded lists are returned in the tiled index tau, like lsseg_get_active. */
lsseg_status lsseg_fwd_record(lsseg_ctx* ctx, const lsseg_params* params, int32_t iters,
                              lsseg_record** out);
/* A record may outlive its context: destroying the context detaches the
   record, which then refuses every call but lsseg_record_destroy. */
lsseg_status lsseg_record_destroy(lsseg_record* rec);
lsseg_status lsseg_record_info(const lsseg_record* rec, int32_t* iters, uint64_t* total_active);
/* Recorded n_k (iters entries), the concatenated lists V_0..V_{iters-1} and one
   decision byte per list entry (total_active entries each; host or device).
   Each V_k holds the voxels of iteration k's active list in the engine's
   spatially coherent traversal order, not in the list's own order. */
lsseg_status lsseg_record_get(const lsseg_record* rec, uint32_t* n_per_iter, uint32_t* lists,
                              uint8_t* decisions);

/* The recorded map at other inputs: both buffers := phi0 (host or device),
   `iters` updates over the recorded lists with the recorded decisions and
   `params`; the final phi goes to phi_out (host or device).  The context's own
   state is not touched. */
lsseg_status lsseg_fwd_replay(lsseg_ctx* ctx, lsseg_record* rec, const float* phi0,
                              const lsseg_params* params, float* phi_out);

/* Adjoint: given dL/dphi of the final phi (host or device), writes dL/dphi0
   (host or device) and dtheta[0..2] = dL/dT, dL/deps, dL/dalpha (host). */
lsseg_status lsseg_bwd(lsseg_ctx* ctx, lsseg_record* rec, const float* dphi, float* dphi0,
                       double* dtheta);

/* ---- measurement aid ----------------------------------------------------

   No counterpart in the reference.  The random-access floor that the path's
   gathers and scatters run against: `reps` launches (one at a time on
   `stream`, each timed by its own start and stop events) in which each of
   `count` threads makes one pseudo-random 4-byte load (mode 0), seven such
   loads (mode 1, the fan-out of the update's stencil without its locality) or
   one such store (mode 2) in the caller's device buffer `buf` of `bytes`
   bytes, each access in its own 32-byte sector.  Mode 2 overwrites words of
   the buffer with zeros.  *us_median receives the median launch time in
   microseconds. */
lsseg_status lsseg_probe_random_access(void* buf, uint64_t bytes, uint32_t count, int32_t mode, int32_t reps,
                                       void* stream, float* us_median);

#ifdef __cplusplus
}
#endif

#endif /* LSSEG_LSSEG_H */
