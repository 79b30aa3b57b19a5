/*
 * lsoracle.c -- CPU restatement of the reference's sparse level-set path.
 * TEST INFRASTRUCTURE ONLY (see lsoracle.h for the contract and layout).
 * Reference paths are relative to code/gpuseg/code/.
 */
#include "lsoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* index helpers                                                              */
/* ------------------------------------------------------------------------- */

/* ComputeIndex3DToTiled1D, Coordinates.cu:65-96 (tile size 16, Config.hpp:92-97) */
uint32_t lso_tau(int32_t X, int32_t Y, int32_t x, int32_t y, int32_t z)
{
    const uint32_t tx = (uint32_t)x >> 4, ty = (uint32_t)y >> 4, tz = (uint32_t)z >> 4;
    const uint32_t TX = (uint32_t)X >> 4, TY = (uint32_t)Y >> 4;
    const uint32_t tile = (tz * TY + ty) * TX + tx;
    const uint32_t within = (((uint32_t)z & 15u) * 16u + ((uint32_t)y & 15u)) * 16u + ((uint32_t)x & 15u);
    return tile * 4096u + within;
}

/* inverse of lso_tau (the reference's ComputeIndexTiled1DTo3D, Coordinates.cu:98-130) */
void lso_untau(int32_t X, int32_t Y, uint32_t t, int32_t* x, int32_t* y, int32_t* z)
{
    const uint32_t TX = (uint32_t)X >> 4, TY = (uint32_t)Y >> 4;
    const uint32_t tile = t >> 12, within = t & 4095u;
    const uint32_t tx = tile % TX, ty = (tile / TX) % TY, tz = tile / (TX * TY);
    *x = (int32_t)(tx * 16u + (within & 15u));
    *y = (int32_t)(ty * 16u + ((within >> 4) & 15u));
    *z = (int32_t)(tz * 16u + (within >> 8));
}

static const int32_t kDir[7][3] = {
    {0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

/* OffsetCoordinates[Unsafe], Coordinates.cu:132-184: an out-of-bounds neighbour
   is replaced by the voxel itself */
uint32_t lso_neighbor(int32_t X, int32_t Y, int32_t Z, uint32_t t, int32_t dir)
{
    int32_t x, y, z;
    lso_untau(X, Y, t, &x, &y, &z);
    const int32_t nx = x + kDir[dir][0], ny = y + kDir[dir][1], nz = z + kDir[dir][2];
    if (nx < 0 || ny < 0 || nz < 0 || nx >= X || ny >= Y || nz >= Z) return t;
    return lso_tau(X, Y, nx, ny, nz);
}

int32_t lso_source_bytes(int32_t dtype)
{
    switch (dtype) {
    case LSO_U8: case LSO_I8: return 1;
    case LSO_U16: case LSO_I16: return 2;
    default: return 4;
    }
}

/* CudaGetWarpAlignedValue, Cuda.cpp:188-200 (alignment 256) */
uint64_t lso_align256(uint64_t n)
{
    return n + ((256u - (n % 256u)) % 256u);
}

/* ShouldOptimizeForFewActiveVoxels, Segmenter.cpp:1363-1369:
   7 * align256(n) * 4 bytes < N * 4 bytes */
int32_t lso_path_is_few(uint64_t n, uint64_t N)
{
    return 7u * lso_align256(n) < N;
}

/* ------------------------------------------------------------------------- */
/* update (float and double instantiations)                                   */
/* ------------------------------------------------------------------------- */
#define LSO_REAL float
#define LSO_FN(name) lso_##name##_f32
#define LSO_SQRT sqrtf
#include "lsoracle_update.inc"
#undef LSO_REAL
#undef LSO_FN
#undef LSO_SQRT

#define LSO_REAL double
#define LSO_FN(name) lso_##name##_f64
#define LSO_SQRT sqrt
#include "lsoracle_update.inc"
#undef LSO_REAL
#undef LSO_FN
#undef LSO_SQRT

/* ------------------------------------------------------------------------- */
/* marking                                                                    */
/* ------------------------------------------------------------------------- */

/* ComputeActiveElementOutputs (UpdateActiveElements.cu:34-103, six-connected,
   temporal culling on) as driven by OutputNewActiveElementsKernel
   (UpdateActiveElementsVolume.cu:304-363), which visits only list positions
   p % 4 in {0,1,2} (the loop `for i < 3` over a uint4 of entries, :329).
   masks[p] bit e is set when the candidate for list e is emitted; bit 3 (self)
   is the OR of the six directional bits.  R is the freshly written buffer. */
void lso_mark(int32_t X, int32_t Y, int32_t Z, const float* R, const uint8_t* flags,
              const uint32_t* V, int64_t n, int32_t mode, uint8_t* masks, int32_t threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t p = 0; p < n; ++p) {
        uint8_t m = 0;
        if ((mode == LSO_MODE_CLEAN || (p & 3) != 3) && flags[p] == 1) {
            const uint32_t t = V[p];
            const float u4 = R[t];
            for (int32_t e = 0; e < 7; ++e) {
                if (e == LSO_SELF) continue;
                const float un = R[lso_neighbor(X, Y, Z, t, e)];
                if (!lso_equals_f32(u4, un, 0.008f)) m |= (uint8_t)(1u << e);
            }
            if (m) m |= (uint8_t)(1u << LSO_SELF);
        }
        masks[p] = m;
    }
}

/* ------------------------------------------------------------------------- */
/* duplicate removal + compaction                                             */
/* ------------------------------------------------------------------------- */

/* "few" path (Segmenter.cpp:883-984):
   - emission into seven lists of stride A = align256(n)
     (UpdateActiveElementsVolume.cu:166-182, 253-301);
   - first-wins duplicate removal with a tag volume: list 0 tags every valid
     entry, lists 1..6 in order drop entries whose tag is set, else set it
     (FilterDuplicates.cu:42-66, 73-193); only positions p < n are scanned;
   - stable compaction of the 7A entries by their valid flags (Cuda.cpp:699-758). */
int64_t lso_compact_few(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                        const uint8_t* masks, uint32_t* out)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    const uint64_t A = lso_align256((uint64_t)n);
    uint32_t* coords = (uint32_t*)calloc(7 * A + 1, sizeof(uint32_t));
    uint8_t* valid = (uint8_t*)calloc(7 * A + 1, 1);
    uint8_t* tag = (uint8_t*)calloc(N, 1);
    for (int64_t p = 0; p < n; ++p) {
        for (int32_t e = 0; e < 7; ++e) {
            if (masks[p] & (1u << e)) {
                coords[(uint64_t)e * A + p] = lso_neighbor(X, Y, Z, V[p], e);
                valid[(uint64_t)e * A + p] = 1;
            }
        }
    }
    for (int64_t p = 0; p < n; ++p)
        if (valid[p]) tag[coords[p]] = 1;
    for (int32_t e = 1; e < 7; ++e) {
        for (int64_t p = 0; p < n; ++p) {
            const uint64_t i = (uint64_t)e * A + p;
            if (valid[i]) {
                if (tag[coords[i]] == 1) valid[i] = 0;
                else tag[coords[i]] = 1;
            }
        }
    }
    int64_t m = 0;
    for (uint64_t i = 0; i < 7 * A; ++i)
        if (valid[i]) out[m++] = coords[i];
    free(coords);
    free(valid);
    free(tag);
    return m;
}

/* "many" path (Segmenter.cpp:985-1019): dense flags at tau(v) for the self bit
   and tau(v+e) for each directional bit (UpdateActiveElementsVolumeKernel,
   UpdateActiveElementsVolume.cu:39-100), then stable compaction of the
   coordinate volume (InitializeCoordinateVolume.cu:37-47) by those flags, which
   yields ascending tau. */
int64_t lso_compact_many(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                         const uint8_t* masks, uint32_t* out)
{
    const uint64_t N = (uint64_t)X * Y * Z;
    uint8_t* flag = (uint8_t*)calloc(N, 1);
    for (int64_t p = 0; p < n; ++p)
        for (int32_t e = 0; e < 7; ++e)
            if (masks[p] & (1u << e)) flag[lso_neighbor(X, Y, Z, V[p], e)] = 1;
    int64_t m = 0;
    for (uint64_t t = 0; t < N; ++t)
        if (flag[t]) out[m++] = (uint32_t)t;
    free(flag);
    return m;
}

/* ------------------------------------------------------------------------- */
/* the same two compactions with OpenMP (the CPU baseline's path)             */
/* ------------------------------------------------------------------------- */

#define LSO_SCAN_BLOCK 4096 /* positions or voxels per counting block */

static int lso_nthreads(int32_t threads) { return threads > 0 ? threads : 1; }

/* Exclusive prefix of cnt[0..k) in place; returns the total. */
static int64_t lso_exclusive_scan(int64_t* cnt, int64_t k)
{
    int64_t run = 0;
    for (int64_t i = 0; i < k; ++i) {
        const int64_t c = cnt[i];
        cnt[i] = run;
        run += c;
    }
    return run;
}

/* lso_compact_few with identical output, parallel over list positions.
   Within one candidate list the targets are distinct (V is duplicate-free and
   a direction shifts voxels injectively; out-of-bounds neighbours are never
   emitted), so first-wins over the list-major concatenation is decided list by
   list in list order (FilterDuplicates.cu:42-66), each list in parallel: a
   candidate is kept iff its target is not yet tagged, and then tags it.  The
   kept candidates are placed by a prefix sum over (list, block of positions).
   tag (N bytes) and keep (n bytes) are caller scratch; tag must be zero on
   entry and is zero again on return (cleared at the placed voxels, which are
   exactly the tagged ones). */
int64_t lso_compact_few_par(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                            const uint8_t* masks, uint32_t* out, uint8_t* tag, uint8_t* keep,
                            int32_t threads)
{
    const int nt = lso_nthreads(threads);
    (void)nt;
    for (int32_t e = 0; e < 7; ++e) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
        for (int64_t p = 0; p < n; ++p) {
            uint8_t k = e ? keep[p] : 0;
            if (masks[p] & (1u << e)) {
                const uint32_t v = lso_neighbor(X, Y, Z, V[p], e);
                if (!tag[v]) {
                    tag[v] = 1;
                    k |= (uint8_t)(1u << e);
                }
            }
            keep[p] = k;
        }
    }
    const int64_t nb = (n + LSO_SCAN_BLOCK - 1) / LSO_SCAN_BLOCK;
    int64_t* cnt = (int64_t*)calloc((size_t)(7 * nb + 1), sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t i = 0; i < 7 * nb; ++i) {
        const int32_t e = (int32_t)(i / nb);
        const int64_t p0 = (i % nb) * LSO_SCAN_BLOCK;
        const int64_t p1 = p0 + LSO_SCAN_BLOCK < n ? p0 + LSO_SCAN_BLOCK : n;
        int64_t c = 0;
        for (int64_t p = p0; p < p1; ++p) c += (keep[p] >> e) & 1u;
        cnt[i] = c;
    }
    const int64_t m = lso_exclusive_scan(cnt, 7 * nb);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t i = 0; i < 7 * nb; ++i) {
        const int32_t e = (int32_t)(i / nb);
        const int64_t p0 = (i % nb) * LSO_SCAN_BLOCK;
        const int64_t p1 = p0 + LSO_SCAN_BLOCK < n ? p0 + LSO_SCAN_BLOCK : n;
        int64_t at = cnt[i];
        for (int64_t p = p0; p < p1; ++p)
            if ((keep[p] >> e) & 1u) out[at++] = lso_neighbor(X, Y, Z, V[p], e);
    }
    free(cnt);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t i = 0; i < m; ++i) tag[out[i]] = 0;
    return m;
}

/* lso_compact_many with identical output: dense flags set in parallel (several
   emitters may set one flag; the stores are atomic), then the flagged voxels
   in ascending tau by a prefix sum over blocks of voxels.  flag (N bytes) is
   caller scratch, zero on entry and on return. */
int64_t lso_compact_many_par(int32_t X, int32_t Y, int32_t Z, const uint32_t* V, int64_t n,
                             const uint8_t* masks, uint32_t* out, uint8_t* flag, int32_t threads)
{
    const int nt = lso_nthreads(threads);
    (void)nt;
    const uint64_t N = (uint64_t)X * Y * Z;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t p = 0; p < n; ++p)
        for (int32_t e = 0; e < 7; ++e)
            if (masks[p] & (1u << e)) __atomic_store_n(&flag[lso_neighbor(X, Y, Z, V[p], e)], 1, __ATOMIC_RELAXED);
    const int64_t nb = (int64_t)((N + LSO_SCAN_BLOCK - 1) / LSO_SCAN_BLOCK);
    int64_t* cnt = (int64_t*)calloc((size_t)(nb + 1), sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t b = 0; b < nb; ++b) {
        const uint64_t t0 = (uint64_t)b * LSO_SCAN_BLOCK;
        const uint64_t t1 = t0 + LSO_SCAN_BLOCK < N ? t0 + LSO_SCAN_BLOCK : N;
        int64_t c = 0;
        for (uint64_t t = t0; t < t1; ++t) c += flag[t] != 0;
        cnt[b] = c;
    }
    const int64_t m = lso_exclusive_scan(cnt, nb);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t b = 0; b < nb; ++b) {
        const uint64_t t0 = (uint64_t)b * LSO_SCAN_BLOCK;
        const uint64_t t1 = t0 + LSO_SCAN_BLOCK < N ? t0 + LSO_SCAN_BLOCK : N;
        int64_t at = cnt[b];
        for (uint64_t t = t0; t < t1; ++t)
            if (flag[t]) {
                out[at++] = (uint32_t)t;
                flag[t] = 0;
            }
    }
    free(cnt);
    return m;
}

/* ------------------------------------------------------------------------- */
/* stateful engine (Segmenter.cpp)                                            */
/* ------------------------------------------------------------------------- */

struct lso_state {
    int32_t nx, ny, nz;      /* unpadded */
    int32_t X, Y, Z;         /* padded to 16 (Segmenter.cpp:1136-1143) */
    uint64_t N;
    int32_t dtype;
    float z_aniso;
    void* src;               /* tiled */
    float* phi[2];           /* phi[rd] is the read buffer */
    int32_t rd;
    int8_t* E;               /* export byte, tiled */
    int8_t* frozen;          /* frozen export (FreezeCurrentSegmentation), tiled */
    uint32_t* V;             /* active list (capacity N) */
    int64_t n;
    uint32_t* Vn;            /* scratch list */
    uint8_t* flags;          /* per position */
    uint8_t* masks;          /* per position */
    int32_t phi_i8;          /* fixed-point phi storage (LEVEL_SET_CHAR, Config.hpp:72-80) */
    uint8_t* scratch;        /* N bytes, zero between iterations: tags / dense flags */
    uint8_t* keep;           /* kept candidate lists per position */
    struct lso_literal* lit; /* the shipped driver's scratch buffers (LITERAL mode), or NULL */
};

/* LITERAL mode: the scratch volumes of the shipped driver, replayed with their
   stale contents (Segmenter.cpp:491-595, 759-1051) */
struct lso_literal {
    uint32_t* coords;        /* coordinate volume; its start doubles as the seven candidate
                                lists of stride align256(n) (Segmenter.cpp:893-897) */
    uint8_t* valid;          /* valid-element volume (0/1; u32 in the reference) */
    uint8_t* tag;            /* tag volume: the update's temporal flags, then the
                                duplicate removal's scratch */
    uint32_t* stamp;         /* per voxel: (iteration, list) of its last candidate, to count
                                duplicates within one list */
    uint32_t stamp_next;
    int32_t coords_initialized; /* mCoordinatesVolumeInitialized */
};

static void* xcalloc(uint64_t count, uint64_t size)
{
    return calloc((size_t)(count ? count : 1), (size_t)size);
}

/* LoadVolumeHostPadded, Segmenter.cpp:1131-1204: pad each axis to a multiple of
   16; padding bytes are 0, or 0x81 (memset of -127) for signed sources.  Float
   sources (an extension) pad with 0.0f.  Stored tiled. */
lso_state* lso_create(int32_t nx, int32_t ny, int32_t nz, int32_t dtype, const void* src_linear,
                      float z_aniso)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || dtype < LSO_U8 || dtype > LSO_F32) return NULL;
    lso_state* s = (lso_state*)calloc(1, sizeof(lso_state));
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->X = (nx + 15) / 16 * 16;
    s->Y = (ny + 15) / 16 * 16;
    s->Z = (nz + 15) / 16 * 16;
    s->N = (uint64_t)s->X * s->Y * s->Z;
    s->dtype = dtype;
    s->z_aniso = z_aniso;
    const int32_t S = lso_source_bytes(dtype);
    const int is_signed = (dtype == LSO_I8 || dtype == LSO_I16 || dtype == LSO_I32);
    s->src = xcalloc(s->N, (uint64_t)S);
    if (is_signed) memset(s->src, 0x81, (size_t)(s->N * (uint64_t)S));
    for (int32_t z = 0; z < nz; ++z)
        for (int32_t y = 0; y < ny; ++y)
            for (int32_t x = 0; x < nx; ++x) {
                const uint64_t li = ((uint64_t)z * ny + y) * nx + x;
                const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                memcpy((char*)s->src + (uint64_t)t * S, (const char*)src_linear + li * S, (size_t)S);
            }
    s->phi[0] = (float*)xcalloc(s->N, sizeof(float));
    s->phi[1] = (float*)xcalloc(s->N, sizeof(float));
    s->E = (int8_t*)xcalloc(s->N, 1);
    s->frozen = (int8_t*)xcalloc(s->N, 1);
    s->V = (uint32_t*)xcalloc(s->N, sizeof(uint32_t));
    s->Vn = (uint32_t*)xcalloc(s->N, sizeof(uint32_t));
    s->flags = (uint8_t*)xcalloc(s->N, 1);
    s->masks = (uint8_t*)xcalloc(s->N, 1);
    s->scratch = (uint8_t*)xcalloc(s->N, 1);
    s->keep = (uint8_t*)xcalloc(s->N, 1);
    /* Segmenter.cpp:150-151 fills both buffers with the int8 value 127 (= 1.0)
       and the export with 127 (:265-266) */
    for (uint64_t i = 0; i < s->N; ++i) {
        s->phi[0][i] = 1.0f;
        s->phi[1][i] = 1.0f;
        s->E[i] = 127;
        s->frozen[i] = 127;
    }
    s->rd = 0;
    s->n = 0;
    return s;
}

void lso_destroy(lso_state* s)
{
    if (!s) return;
    free(s->src);
    free(s->phi[0]);
    free(s->phi[1]);
    free(s->E);
    free(s->frozen);
    free(s->V);
    free(s->Vn);
    free(s->flags);
    free(s->masks);
    free(s->scratch);
    free(s->keep);
    if (s->lit) {
        free(s->lit->coords);
        free(s->lit->valid);
        free(s->lit->tag);
        free(s->lit->stamp);
        free(s->lit);
    }
    free(s);
}

void lso_dims(const lso_state* s, int32_t* X, int32_t* Y, int32_t* Z)
{
    *X = s->X; *Y = s->Y; *Z = s->Z;
}

static int8_t export_byte(float v)
{
    /* __float2int_rd(v * 127.0f) (UpdateLevelSetVolume.cu:323), saturated to the
       int8 range for injected values outside [-1, 1] */
    double q = floor((double)(v * 127.0f));
    if (q > 127.0) q = 127.0;
    if (q < -128.0) q = -128.0;
    return (int8_t)(int32_t)q;
}

/* The value a stored phi reads back as.  FP32 storage: itself.  Fixed point
   (LEVEL_SET_CHAR, Config.hpp:72-80): stored as __float2int_rd(v * 127)
   (UpdateLevelSetVolume.cu:316-318), read through a normalised-float texture
   as s / 127 with -128 -> -1 (Textures.cu:12). */
static float phi_stored(const lso_state* s, float v)
{
    if (!s->phi_i8) return v;
    const int8_t q = export_byte(v);
    return q == -128 ? -1.0f : (float)q / 127.0f;
}

void lso_set_phi_storage(lso_state* s, int32_t fixed_point)
{
    s->phi_i8 = fixed_point ? 1 : 0;
}

/* fixed point: the values just written at the list voxels become what they read back as */
static void store_list(lso_state* s, float* W)
{
    if (!s->phi_i8) return;
    for (int64_t p = 0; p < s->n; ++p) W[s->V[p]] = phi_stored(s, W[s->V[p]]);
}

/* Build extension used by the benchmark configurations: both buffers hold phi0,
   the export byte is floor(127 phi0), and V0 = {|phi0| < band} in ascending tau
   (the order a dense compaction produces, Cuda.cpp:699-758). */
int32_t lso_set_phi(lso_state* s, const float* phi0_linear, float band)
{
    for (uint64_t i = 0; i < s->N; ++i) {
        s->phi[0][i] = 1.0f;
        s->E[i] = 127;
    }
    for (int32_t z = 0; z < s->nz; ++z)
        for (int32_t y = 0; y < s->ny; ++y)
            for (int32_t x = 0; x < s->nx; ++x) {
                const uint64_t li = ((uint64_t)z * s->ny + y) * s->nx + x;
                const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                s->phi[0][t] = phi0_linear[li];
                s->E[t] = export_byte(phi0_linear[li]);
            }
    s->rd = 0;
    int64_t m = 0;
    for (uint64_t t = 0; t < s->N; ++t) {
        const float v = s->phi[0][t];
        if ((v < 0 ? -v : v) < band) s->V[m++] = (uint32_t)t;
        s->phi[0][t] = phi_stored(s, v); /* the band uses phi0 itself */
    }
    memcpy(s->phi[1], s->phi[0], (size_t)(s->N * sizeof(float)));
    s->n = m;
    return 0;
}

/* Segmenter::InitializeSeed (Segmenter.cpp:491-595): two passes of
   InitializeLevelSetKernel (InitializeLevelSetVolume.cu:55-98) with a buffer swap
   after each, then dense keep-flags from the spatial gradient
   (InitializeActiveElementsVolume.cu:37-74) compacted in ascending tau. */
int32_t lso_seed_sphere(lso_state* s, float cx, float cy, float cz, uint32_t radius)
{
    const uint32_t sx = (uint32_t)cx, sy = (uint32_t)cy, sz = (uint32_t)cz; /* :502-505 */
    const int32_t r = (int32_t)radius;
    const float smooth = 4.0f; /* Config.hpp:61 */
    for (int pass = 0; pass < 2; ++pass) {
        const float* R = s->phi[s->rd];
        float* W = s->phi[1 - s->rd];
        for (int32_t z = 0; z < s->Z; ++z)
            for (int32_t y = 0; y < s->Y; ++y)
                for (int32_t x = 0; x < s->X; ++x) {
                    const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                    const int32_t vx = (int32_t)(sx - (uint32_t)x);
                    const int32_t vy = (int32_t)(sy - (uint32_t)y);
                    const float vz = ((float)sz - (float)z) * s->z_aniso;
                    const float d2 = (float)vx * (float)vx + (float)vy * (float)vy + vz * vz;
                    const float d = sqrtf(d2);
                    float c = 1.0f;
                    if (d < (float)r) c = -1.0f;
                    else if (d < (float)r + smooth) c = -1.0f + ((d - (float)r) / (smooth / 2.0f));
                    float nv = R[t] + c - 1.0f;
                    if (nv < -1.0f) nv = -1.0f;
                    if (nv > 1.0f) nv = 1.0f;
                    W[t] = phi_stored(s, nv);
                    s->E[t] = export_byte(nv);
                }
        s->rd = 1 - s->rd; /* SwapLevelSetBuffers, Segmenter.cpp:1339-1351 */
    }
    lso_reinit_active(s);
    return 0;
}

/* dense keep-flags from the spatial gradient of the read buffer
   (InitializeActiveElementsVolumeKernel, InitializeActiveElementsVolume.cu:37-74)
   compacted in ascending tau (Segmenter.cpp:536-554, 847-876) */
void lso_reinit_active(lso_state* s)
{
    const float* R = s->phi[s->rd];
    int64_t m = 0;
    for (uint64_t t = 0; t < s->N; ++t) {
        const float u4 = R[t];
        int flat = 1;
        for (int32_t e = 0; e < 7 && flat; ++e) {
            if (e == LSO_SELF) continue;
            if (!lso_equals_f32(u4, R[lso_neighbor(s->X, s->Y, s->Z, (uint32_t)t, e)], 0.001f)) flat = 0;
        }
        if (!flat) s->V[m++] = (uint32_t)t;
    }
    s->n = m;
}

/* the parameter-change sub-step of Segmenter::RequestUpdateSegmentation
   (Segmenter.cpp:811-876): update, swap, restart the active set */
void lso_iterate_reinit(lso_state* s, float T, float eps, float alpha, float dt)
{
    lso_update_f32(s->X, s->Y, s->Z, s->dtype, s->src, s->phi[s->rd], s->phi[1 - s->rd], s->E, s->flags,
                   s->V, s->n, T, eps, alpha, dt, 1);
    store_list(s, s->phi[1 - s->rd]);
    s->rd = 1 - s->rd;
    lso_reinit_active(s);
}

/* One sub-step of Segmenter::RequestUpdateSegmentation (Segmenter.cpp:772-1036):
   update R -> W, swap, mark on the new read buffer, then the few or many path. */
int32_t lso_iterate(lso_state* s, float T, float eps, float alpha, float dt, int32_t mode,
                    int32_t iters, int32_t threads,
                    uint32_t* n_hist, uint64_t* c_hist, uint32_t* m_hist)
{
    for (int32_t it = 0; it < iters; ++it) {
        const float* R = s->phi[s->rd];
        float* W = s->phi[1 - s->rd];
        lso_update_f32(s->X, s->Y, s->Z, s->dtype, s->src, R, W, s->E, s->flags, s->V, s->n,
                       T, eps, alpha, dt, threads);
        store_list(s, W);
        s->rd = 1 - s->rd;
        lso_mark(s->X, s->Y, s->Z, s->phi[s->rd], s->flags, s->V, s->n, mode, s->masks, threads);
        uint64_t c = 0;
        const uint8_t* masks = s->masks;
        const int64_t n = s->n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : c) num_threads(threads > 0 ? threads : 1)
#endif
        for (int64_t p = 0; p < n; ++p) c += (uint64_t)__builtin_popcount(masks[p]);
        /* the compactions of lso_compact_few / lso_compact_many, in parallel
           with identical output (tests/test_oracle_lists.py pins them) */
        int64_t m;
        if (lso_path_is_few((uint64_t)s->n, s->N))
            m = lso_compact_few_par(s->X, s->Y, s->Z, s->V, s->n, s->masks, s->Vn, s->scratch, s->keep, threads);
        else
            m = lso_compact_many_par(s->X, s->Y, s->Z, s->V, s->n, s->masks, s->Vn, s->scratch, threads);
        if (n_hist) n_hist[it] = (uint32_t)s->n;
        if (c_hist) c_hist[it] = c;
        if (m_hist) m_hist[it] = (uint32_t)m;
        uint32_t* tmp = s->V;
        s->V = s->Vn;
        s->Vn = tmp;
        s->n = m;
    }
    return 0;
}

int64_t lso_active_count(const lso_state* s) { return s->n; }

int64_t lso_get_active(const lso_state* s, uint32_t* out, int64_t cap)
{
    const int64_t k = s->n < cap ? s->n : cap;
    memcpy(out, s->V, (size_t)k * sizeof(uint32_t));
    return s->n;
}

void lso_get_state(const lso_state* s, float* phi_read, float* phi_write, int8_t* E)
{
    if (phi_read) memcpy(phi_read, s->phi[s->rd], (size_t)(s->N * sizeof(float)));
    if (phi_write) memcpy(phi_write, s->phi[1 - s->rd], (size_t)(s->N * sizeof(float)));
    if (E) memcpy(E, s->E, (size_t)s->N);
}

int32_t lso_set_state(lso_state* s, const float* phi_read, const float* phi_write,
                      const int8_t* E, const uint32_t* V, int64_t n)
{
    if (n < 0 || (uint64_t)n > s->N) return -1;
    for (int64_t p = 0; p < n; ++p)
        if ((uint64_t)V[p] >= s->N) return -1;
    s->rd = 0;
    memcpy(s->phi[0], phi_read, (size_t)(s->N * sizeof(float)));
    memcpy(s->phi[1], phi_write, (size_t)(s->N * sizeof(float)));
    memcpy(s->E, E, (size_t)s->N);
    memcpy(s->V, V, (size_t)n * sizeof(uint32_t));
    s->n = n;
    return 0;
}

void lso_get_phi(const lso_state* s, int32_t which, float* out_linear)
{
    const float* P = s->phi[which == 0 ? s->rd : 1 - s->rd];
    for (int32_t z = 0; z < s->nz; ++z)
        for (int32_t y = 0; y < s->ny; ++y)
            for (int32_t x = 0; x < s->nx; ++x)
                out_linear[((uint64_t)z * s->ny + y) * s->nx + x] = P[lso_tau(s->X, s->Y, x, y, z)];
}

static int8_t add_export_byte(int8_t ra, int8_t rb)
{
    /* AddLevelSetKernel, InitializeLevelSetVolume.cu:160-168 */
    const float a = (float)ra / 127.0f;
    const float b = (float)rb / 127.0f;
    float r = a + b - 1.0f;
    if (r < -1.0f) r = -1.0f;
    if (r > 1.0f) r = 1.0f;
    return (int8_t)(int32_t)floor((double)(r * 127.0f));
}

/* GetSaveableSegmentation + UnpadVolume (Segmenter.cpp:1216-1337): the export is
   copied, unioned with the frozen export by AddLevelSetKernel (:1231-1234; the
   frozen export starts at 127, :265-266) and thresholded at <= 0 (:1251-1262),
   over the unpadded box */
void lso_get_mask(const lso_state* s, uint8_t* out_linear)
{
    for (int32_t z = 0; z < s->nz; ++z)
        for (int32_t y = 0; y < s->ny; ++y)
            for (int32_t x = 0; x < s->nx; ++x) {
                const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                out_linear[((uint64_t)z * s->ny + y) * s->nx + x] =
                    add_export_byte(s->E[t], s->frozen[t]) <= 0 ? 1 : 0;
            }
}

/* a source voxel as the reference reads it for the statistics: static_cast<float> */
static float source_value(const lso_state* s, uint32_t t)
{
    switch (s->dtype) {
    case LSO_U8: return (float)((const uint8_t*)s->src)[t];
    case LSO_I8: return (float)((const int8_t*)s->src)[t];
    case LSO_U16: return (float)((const uint16_t*)s->src)[t];
    case LSO_I16: return (float)((const int16_t*)s->src)[t];
    case LSO_U32: return (float)((const uint32_t*)s->src)[t];
    case LSO_I32: return (float)((const int32_t*)s->src)[t];
    default: return ((const float*)s->src)[t];
    }
}

/* FreezeCurrentSegmentation (Segmenter.cpp:659-679): frozen := union(frozen, export) */
void lso_freeze(lso_state* s)
{
    for (uint64_t t = 0; t < s->N; ++t) s->frozen[t] = add_export_byte(s->frozen[t], s->E[t]);
}

/* CollectCurrentLevelSetVoxels + CalculateSegmentationParameters
   (Segmenter.cpp:357-478): in the reference's order (the padded volume, x
   fastest) and arithmetic (sequential float sums), the mean and the sample
   standard deviation of the source over the voxels whose export byte is <= 0.
   Returns the number of such voxels. */
int64_t lso_auto_parameters(const lso_state* s, float* T, float* eps)
{
    float total = 0.0f;
    int64_t n = 0;
    for (int32_t z = 0; z < s->Z; ++z)
        for (int32_t y = 0; y < s->Y; ++y)
            for (int32_t x = 0; x < s->X; ++x) {
                const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                if (s->E[t] <= 0) {
                    total += source_value(s, t);
                    ++n;
                }
            }
    const float count = (float)n;
    const float mean = total / count;
    float sd = 0.0f;
    for (int32_t z = 0; z < s->Z; ++z)
        for (int32_t y = 0; y < s->Y; ++y)
            for (int32_t x = 0; x < s->X; ++x) {
                const uint32_t t = lso_tau(s->X, s->Y, x, y, z);
                if (s->E[t] <= 0) {
                    const float d = source_value(s, t) - mean;
                    sd += d * d;
                }
            }
    *T = mean;
    *eps = (float)sqrt(sd / (count - 1.0f));
    return n;
}

/* ClearCurrentSegmentation (Segmenter.cpp:597-640): phi := 1, export := 127,
   empty list; all != 0 also clears the frozen export (ClearAllSegmentations,
   :641-657) */
void lso_clear(lso_state* s, int32_t all)
{
    for (uint64_t t = 0; t < s->N; ++t) {
        s->phi[0][t] = 1.0f;
        s->phi[1][t] = 1.0f;
        s->E[t] = 127;
        if (all) s->frozen[t] = 127;
    }
    s->n = 0;
}

/* AddLevelSetKernel, InitializeLevelSetVolume.cu:151-169 */
void lso_add_export(int8_t* result, const int8_t* add, int64_t N)
{
    for (int64_t i = 0; i < N; ++i) result[i] = add_export_byte(result[i], add[i]);
}

/* ------------------------------------------------------------------------- */
/* LITERAL mode (diagnostic): the shipped driver's buffer artefacts           */
/* ------------------------------------------------------------------------- */
/* REF mode above gives every iteration clean scratch.  The shipped driver
   reuses three volumes across iterations, and three artefacts follow:
   1. the dense keep-flags that seed initialisation writes into the valid
      volume (Segmenter.cpp:536-554) are never cleared; the few path clears the
      valid volume only over [0, 7 align256(n)) after each compaction (:977),
      so a flag at tau j is compacted once 7 align256(n) exceeds j, paired with
      whatever the coordinate volume holds at j, and it escapes duplicate
      removal, which scans only positions p < n of each list
      (FilterDuplicates.cu:90-100, 160-170);
   2. the many path never clears the tag volume (:985-1019), which holds the
      update's temporal flags (UpdateLevelSetVolume.cu:327-331); after a switch
      back to the few path, a stale flag of 1 at a voxel outside the current
      list makes duplicate removal drop candidates of lists 1..6 for it
      (FilterDuplicates.cu:127-153);
   3. duplicates in a list (from 1) break the uniqueness within each candidate
      list that duplicate removal relies on (main.tex:399): the reference's
      read-then-tag kernel then races; this replay resolves it in position
      order and counts it.
   lso_literal_begin sets the buffers as the reference's initialisation leaves
   them for the current list (coordinate volume initialised, the list's voxels
   flagged in the valid volume, tags clear).  On the dense path the stale flags
   of 1 are all compacted at once: a first dense iteration re-activates the
   whole initial list.  lso_iterate_literal replays
   iterations with them, in REF mode's marking (the p % 4 == 3 skip). */

int32_t lso_literal_begin(lso_state* s, int32_t initial_flags)
{
    if (!s->lit) {
        s->lit = (struct lso_literal*)calloc(1, sizeof(struct lso_literal));
        if (!s->lit) return -1;
        s->lit->coords = (uint32_t*)xcalloc(s->N, sizeof(uint32_t));
        s->lit->valid = (uint8_t*)xcalloc(s->N, 1);
        s->lit->tag = (uint8_t*)xcalloc(s->N, 1);
        s->lit->stamp = (uint32_t*)xcalloc(s->N, sizeof(uint32_t));
        if (!s->lit->coords || !s->lit->valid || !s->lit->tag || !s->lit->stamp) return -1;
    }
    struct lso_literal* L = s->lit;
    /* CudaInitializeCoordinateVolume (InitializeCoordinateVolume.cu:37-47): the
       entry at tau holds the voxel tau */
    for (uint64_t t = 0; t < s->N; ++t) {
        L->coords[t] = (uint32_t)t;
        L->valid[t] = 0;
        L->tag[t] = 0;
        L->stamp[t] = 0;
    }
    L->stamp_next = 1;
    L->coords_initialized = 1;
    /* the dense keep-flags whose compaction gave the current list (initial_flags
       = 0 starts with a clean valid volume instead, to isolate the other two
       artefacts) */
    if (initial_flags)
        for (int64_t p = 0; p < s->n; ++p) L->valid[s->V[p]] = 1;
    return 0;
}

/* events[3 * it + 0]: stale valid flags compacted (artefact 1),
   events[3 * it + 1]: candidates of lists 1..6 whose target carried a stale
   tag at the start of duplicate removal (artefact 2),
   events[3 * it + 2]: valid entries repeating a target within one list (3) */
int32_t lso_iterate_literal(lso_state* s, float T, float eps, float alpha, float dt, int32_t iters,
                            int32_t threads, uint32_t* n_hist, uint64_t* c_hist, uint32_t* m_hist,
                            uint64_t* events)
{
    struct lso_literal* L = s->lit;
    if (!L) return -1;
    for (int32_t it = 0; it < iters; ++it) {
        const int64_t n = s->n;
        const float* R = s->phi[s->rd];
        float* W = s->phi[1 - s->rd];
        lso_update_f32(s->X, s->Y, s->Z, s->dtype, s->src, R, W, s->E, s->flags, s->V, n, T, eps, alpha, dt,
                       threads);
        store_list(s, W);
        /* the update stores the temporal flag in the tag volume, where the
           marking reads it (UpdateActiveElements.cu:75-103) */
        for (int64_t p = 0; p < n; ++p) L->tag[s->V[p]] = s->flags[p];
        for (int64_t p = 0; p < n; ++p) s->flags[p] = L->tag[s->V[p]];
        s->rd = 1 - s->rd;
        lso_mark(s->X, s->Y, s->Z, s->phi[s->rd], s->flags, s->V, n, LSO_MODE_REF, s->masks, threads);
        uint64_t c = 0, ev_stale = 0, ev_tag = 0, ev_dup = 0;
        for (int64_t p = 0; p < n; ++p) c += (uint64_t)__builtin_popcount(s->masks[p]);
        int64_t m = 0;
        if (lso_path_is_few((uint64_t)n, s->N)) {
            const uint64_t A = lso_align256((uint64_t)n);
            for (uint64_t i = 0; i < 7 * A; ++i) ev_stale += L->valid[i];
            /* CudaOutputNewActiveElements (UpdateActiveElementsVolume.cu:129-363):
               coordinates and valid flags of the emitted slots only */
            for (int64_t p = 0; p < n; ++p)
                for (int32_t e = 0; e < 7; ++e)
                    if (s->masks[p] & (1u << e)) {
                        L->coords[(uint64_t)e * A + p] = lso_neighbor(s->X, s->Y, s->Z, s->V[p], e);
                        L->valid[(uint64_t)e * A + p] = 1;
                    }
            L->coords_initialized = 0; /* UpdateHostStateOptimizeForFewActiveVoxels, :1511-1514 */
            /* CudaMemSetCharSparse(tag, 0, n) at the current list (:933) */
            for (int64_t p = 0; p < n; ++p) L->tag[s->V[p]] = 0;
            /* duplicates within a list, and targets tagged before the filter runs */
            for (int32_t e = 0; e < 7; ++e) {
                const uint32_t st = L->stamp_next++;
                for (int64_t p = 0; p < n; ++p) {
                    const uint64_t i = (uint64_t)e * A + (uint64_t)p;
                    if (!L->valid[i]) continue;
                    const uint32_t v = L->coords[i];
                    if (L->stamp[v] == st) ++ev_dup;
                    L->stamp[v] = st;
                    if (e > 0 && L->tag[v]) ++ev_tag;
                }
            }
            /* CudaFilterDuplicates (FilterDuplicates.cu:16-193): positions p < n */
            for (int64_t p = 0; p < n; ++p)
                if (L->valid[p]) L->tag[L->coords[p]] = 1;
            for (int32_t e = 1; e < 7; ++e)
                for (int64_t p = 0; p < n; ++p) {
                    const uint64_t i = (uint64_t)e * A + (uint64_t)p;
                    if (!L->valid[i]) continue;
                    if (L->tag[L->coords[i]] == 1) L->valid[i] = 0;
                    else L->tag[L->coords[i]] = 1;
                }
            /* CudppCompact over 7 align256(n) slots (:963-970), then
               CudaMemSet(valid, 0, 7 align256(n)) (:977) */
            for (uint64_t i = 0; i < 7 * A; ++i)
                if (L->valid[i]) {
                    s->Vn[m++] = L->coords[i];
                    L->valid[i] = 0;
                }
        } else {
            if (!L->coords_initialized) { /* :992-996 */
                for (uint64_t t = 0; t < s->N; ++t) L->coords[t] = (uint32_t)t;
                L->coords_initialized = 1;
            }
            for (uint64_t t = 0; t < s->N; ++t) ev_stale += L->valid[t];
            /* CudaUpdateActiveElementsVolume (UpdateActiveElementsVolume.cu:8-100) */
            for (int64_t p = 0; p < n; ++p)
                for (int32_t e = 0; e < 7; ++e)
                    if (s->masks[p] & (1u << e)) L->valid[lso_neighbor(s->X, s->Y, s->Z, s->V[p], e)] = 1;
            /* CudppCompact over the volume (:1009-1016), then
               CudaMemSetIntSparse(valid, 0, n') at the new list (:1018) */
            for (uint64_t t = 0; t < s->N; ++t)
                if (L->valid[t]) s->Vn[m++] = L->coords[t];
            for (int64_t p = 0; p < m; ++p) L->valid[s->Vn[p]] = 0;
        }
        if (L->stamp_next > 0xFFFFFF00u) { /* restamp before the counter wraps */
            memset(L->stamp, 0, (size_t)(s->N * sizeof(uint32_t)));
            L->stamp_next = 1;
        }
        if (n_hist) n_hist[it] = (uint32_t)n;
        if (c_hist) c_hist[it] = c;
        if (m_hist) m_hist[it] = (uint32_t)m;
        if (events) {
            events[3 * it + 0] = ev_stale;
            events[3 * it + 1] = ev_tag;
            events[3 * it + 2] = ev_dup;
        }
        uint32_t* tmp = s->V;
        s->V = s->Vn;
        s->Vn = tmp;
        s->n = m;
    }
    return 0;
}
