// lsseg.hip -- sparse level-set segmentation engine for MI355X (gfx950).
//
// Implements the C ABI of include/lsseg/lsseg.h.  Reference paths in comments
// are relative to code/gpuseg/code/ of the reference repository.
//
// Per iteration (parity q: read phi[q], list[q]; write phi[q^1], list[q^1]):
//   k_update   one lane per list position: 7-point stencil on phi[q], source
//              sample, new phi, export byte, time-derivative flag
//              (UpdateLevelSetVolume.cu:222-342).  Also clears the previous
//              iteration's per-voxel emission masks.
//   k_mark     one lane per list position: emission mask over the seven
//              candidate lists (UpdateActiveElements.cu:75-103) with the
//              p % 4 == 3 skip of the reference (UpdateActiveElementsVolume.cu:329);
//              sparse path stores masks per voxel, dense path sets voxel flags.
//   k_compact  duplicate removal + stable compaction in one pass: sparse path
//              scans the 7n candidate slots in list-major order and keeps a
//              candidate unless a generator in an earlier list emits the same
//              voxel (first-wins, FilterDuplicates.cu:42-66); dense path scans
//              voxel flags in tiled order (Segmenter.cpp:985-1019).  Offsets come
//              from a single-pass decoupled look-back scan (Merrill & Garland,
//              "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016).
// All counts stay in device memory; the host never reads n between iterations.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lsseg/lsseg.h"

namespace lsseg {

// ---------------------------------------------------------------------------
// constants (Config.hpp:55-66, 92-97)
// ---------------------------------------------------------------------------
constexpr float kTemporalThreshold = 0.008f;  // TEMPORAL_DERIVATIVE_THRESHOLD
constexpr float kSpatialThreshold = 0.001f;   // SPATIAL_DERIVATIVE_THRESHOLD
constexpr float kRescale = 0.01f;             // LEVEL_SET_RESCALE_AMOUNT
constexpr float kSmoothDistance = 4.0f;       // LEVEL_SET_INITIALIZE_SMOOTH_DISTANCE
constexpr float kConnTol = 0.1f;              // UpdateLevelSetVolume.cu:282-287

constexpr int kBlock = 256;                   // threads per block (4 waves of 64)
constexpr int kItems = 16;                    // compaction slots per thread
constexpr int kScanTile = kBlock * kItems;    // 4096 slots per look-back tile
constexpr uint32_t kSpinLimit = 1u << 24;     // bounded look-back wait

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;

enum : uint32_t { kErrSpin = 1u };

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct FastDiv {
    uint32_t d, m, s;
};

static FastDiv make_fastdiv(uint32_t d)
{
    // q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = 2^32 (2^s - d) / d + 1
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    f.s = s;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
    return f;
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f)
{
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t hi = __umulhi(n, f.m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi + n) >> f.s);
}

struct Geom {
    uint32_t X, Y, Z;        // padded dims
    uint32_t TX, TY, TZ;     // tiles per axis
    uint32_t sTY, sTZ;       // tiled strides of one tile row (y) and tile slab (z)
    uint64_t N;
    FastDiv dTX, dTXY;
};

// Out-of-bounds neighbours map to the voxel itself (Coordinates.cu:157-184);
// bit e of *inb is set when neighbour e (list order) is inside the volume.
__device__ inline void neighbors(const Geom& g, uint32_t t, uint32_t nb[7], uint32_t* inb)
{
    const uint32_t lx = t & 15u, ly = (t >> 4) & 15u, lz = (t >> 8) & 15u;
    const uint32_t tile = t >> 12;
    const uint32_t qy = fdiv(tile, g.dTX);
    const uint32_t tx = tile - qy * g.TX;
    const uint32_t tz = fdiv(tile, g.dTXY);
    const uint32_t ty = qy - tz * g.TY;
    uint32_t in = 0;
    nb[0] = lz ? t - 256u : (tz ? t - g.sTZ + 3840u : t);
    in |= (lz | tz) ? 1u : 0u;
    nb[1] = ly ? t - 16u : (ty ? t - g.sTY + 240u : t);
    in |= (ly | ty) ? 2u : 0u;
    nb[2] = lx ? t - 1u : (tx ? t - 4096u + 15u : t);
    in |= (lx | tx) ? 4u : 0u;
    nb[3] = t;
    nb[4] = lx < 15u ? t + 1u : (tx + 1u < g.TX ? t + 4096u - 15u : t);
    in |= (lx < 15u || tx + 1u < g.TX) ? 16u : 0u;
    nb[5] = ly < 15u ? t + 16u : (ty + 1u < g.TY ? t + g.sTY - 240u : t);
    in |= (ly < 15u || ty + 1u < g.TY) ? 32u : 0u;
    nb[6] = lz < 15u ? t + 256u : (tz + 1u < g.TZ ? t + g.sTZ - 3840u : t);
    in |= (lz < 15u || tz + 1u < g.TZ) ? 64u : 0u;
    *inb = in;
}

__device__ inline void decode_xyz(const Geom& g, uint32_t t, uint32_t* x, uint32_t* y, uint32_t* z)
{
    const uint32_t tile = t >> 12;
    const uint32_t qy = fdiv(tile, g.dTX);
    const uint32_t tx = tile - qy * g.TX;
    const uint32_t tz = fdiv(tile, g.dTXY);
    const uint32_t ty = qy - tz * g.TY;
    *x = tx * 16u + (t & 15u);
    *y = ty * 16u + ((t >> 4) & 15u);
    *z = tz * 16u + ((t >> 8) & 15u);
}

__host__ __device__ inline uint64_t align256(uint64_t n) { return n + ((256u - (n % 256u)) % 256u); }

// ShouldOptimizeForFewActiveVoxels (Segmenter.cpp:1363-1369)
__host__ __device__ inline bool path_is_few(uint64_t n, uint64_t N) { return 7u * align256(n) < N; }

// ---------------------------------------------------------------------------
// device-resident control block and parameters
// ---------------------------------------------------------------------------
struct DevCtl {
    uint32_t n[2];        // entries in list[0], list[1]
    uint32_t ticket;      // look-back tile ticket
    uint32_t iter;        // iteration index within the current lsseg_iterate call
    uint32_t stats_cap;
    uint32_t error;
    uint64_t c_acc;       // candidate emissions of the current iteration
    uint32_t few_iters, many_iters;
};

struct DevParams {
    float negInv;         // -1 / eps  (eps truncated to an integer for int sources)
    float Tf;             // T for float sources
    int32_t Ti;           // (int)T for int sources
    float oneMinusAlpha;  // 1 - alpha
    float alpha;
    float dtAbs;          // |dt|
    int32_t flipDensity;  // dt < 0 (UpdateLevelSetVolume.cu:297-301)
    int32_t alphaPositive;
    int32_t refMode;      // skip positions p % 4 == 3
    int32_t pad_;
};

struct KArgs {
    Geom g;
    float* phi[2];
    const void* src;
    int8_t* E;
    int8_t* frozen;
    uint8_t* M;           // per-voxel emission masks (sparse path)
    uint8_t* F;           // per-voxel flags (dense path, initialisation)
    uint32_t* list[2];
    uint8_t* fl;          // per-position time-derivative flags
    uint8_t* mlist;       // per-position emission masks
    uint64_t* tstate;     // look-back tile states
    DevCtl* ctl;
    const DevParams* prm;
    uint32_t* n_hist;
    uint64_t* c_hist;
    uint32_t* m_hist;
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline uint64_t lanemask_lt(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

__device__ inline uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline uint32_t wave_inclusive_scan(uint32_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ inline uint64_t ld_state(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_state(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back for tile `tile` with aggregate `total`; executed by all 64
// lanes of one wave.  The tile's count travels inside the 64-bit status word,
// so no other memory needs to be ordered with it.  Returns the exclusive prefix.
__device__ uint32_t lookback(uint64_t* states, uint32_t tile, uint32_t total, uint32_t* err)
{
    const int lane = threadIdx.x & 63;
    if (tile == 0) {
        if (lane == 0) st_state(&states[0], kFlagPrefix | total);
        return 0;
    }
    if (lane == 0) st_state(&states[tile], kFlagAgg | total);
    uint32_t excl = 0;
    int64_t look = (int64_t)tile - 1;
    uint32_t spins = 0;
    for (;;) {
        const int64_t idx = look - lane;
        uint64_t s = idx >= 0 ? ld_state(&states[idx]) : kFlagPrefix;
        for (;;) {
            const bool pending = (s >> 62) == 0;
            if (__ballot(pending) == 0) break;
            if (++spins > kSpinLimit) {
                if (lane == 0) atomicOr(err, kErrSpin);
                s = kFlagPrefix;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            if (pending) s = ld_state(&states[idx]);
        }
        const uint64_t pm = __ballot((s >> 62) == 2);
        uint32_t v = (uint32_t)s;
        if (pm) {
            const int k = __ffsll((unsigned long long)pm) - 1;
            excl += wave_sum(lane <= k ? v : 0u);
            break;
        }
        excl += wave_sum(v);
        look -= 64;
    }
    if (lane == 0) st_state(&states[tile], kFlagPrefix | (uint64_t)(excl + total));
    return excl;
}

template <typename T>
__device__ inline int32_t source_int(const void* src, uint32_t t)
{
    return (int32_t)(static_cast<const T*>(src)[t]);
}

__device__ inline float fmin_ref(float a, float b) { return b < a ? b : a; }
__device__ inline float fmax_ref(float a, float b) { return b > a ? b : a; }

__device__ inline int8_t export_byte(float v)
{
    // __float2int_rd(v * 127) (UpdateLevelSetVolume.cu:323); saturated for
    // injected values outside [-1, 1]
    float q = floorf(v * 127.0f);
    q = q > 127.0f ? 127.0f : (q < -128.0f ? -128.0f : q);
    return (int8_t)(int32_t)q;
}

// ---------------------------------------------------------------------------
// k_update
// ---------------------------------------------------------------------------
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_update(KArgs a, int q)
{
    const uint32_t n = a.ctl->n[q];
    const uint32_t nprev = a.ctl->n[q ^ 1];
    const bool clear_prev = path_is_few(nprev, a.g.N);
    const uint32_t lim = n > nprev ? n : nprev;
    const uint32_t* __restrict__ V = a.list[q];
    const uint32_t* __restrict__ Vprev = a.list[q ^ 1];
    const float* __restrict__ R = a.phi[q];
    float* __restrict__ W = a.phi[q ^ 1];
    const DevParams P = *a.prm;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < lim; p += stride) {
        // the previous iteration's masks (still indexed by its positions)
        if (clear_prev && p < nprev && a.mlist[p]) a.M[Vprev[p]] = 0;
        if (p >= n) continue;
        const uint32_t t = V[p];
        uint32_t nb[7], inb;
        neighbors(a.g, t, nb, &inb);
        const float unz = R[nb[0]];
        const float u1 = R[nb[1]];
        const float u3 = R[nb[2]];
        const float u4 = R[t];
        const float u5 = R[nb[4]];
        const float u7 = R[nb[5]];
        const float upz = R[nb[6]];

        const float dcx = (u5 - u3) / 2.0f;
        const float dcy = (u7 - u1) / 2.0f;
        const float dcz = (upz - unz) / 2.0f;
        const float fx = u5 - u4, fy = u7 - u4, fz = upz - u4;
        const float bx = u4 - u3, by = u4 - u1, bz = u4 - unz;
        const float sdx = fx - bx, sdy = fy - by, sdz = fz - bz;

        float conn = 0.0f;
        if (P.alphaPositive) {
            if (fabsf(upz - unz) <= kConnTol && fabsf(unz - u1) <= kConnTol &&
                fabsf(u1 - u3) <= kConnTol && fabsf(u3 - u5) <= kConnTol &&
                fabsf(u5 - u7) <= kConnTol && fabsf(u7 - upz) <= kConnTol) {
                const float avg = (upz + u1 + u3 + u5 + u7 + unz) / 6.0f;
                conn = avg - u4;
            }
        }
        const float curv = sdx + sdy + sdz;
        float f;
        if constexpr (kFloatSrc) {
            const float s = static_cast<const float*>(a.src)[t];
            const float d = s - P.Tf;
            f = P.negInv * fabsf(d) + 1.0f;
        } else {
            const int32_t s = source_int<SrcT>(a.src, t);
            const int32_t d = s - P.Ti;
            const int32_t ad = d < 0 ? -d : d;
            f = P.negInv * (float)ad + 1.0f;
        }
        const float dens = P.flipDensity ? f : -f;
        const float speed = (P.oneMinusAlpha * dens) + (P.alpha * curv);
        const float g = __builtin_sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
        const float delta = ((P.dtAbs * speed * g)) + conn;
        float nv = fmin_ref(1.0f, fmax_ref(-1.0f, u4 + delta));
        nv = fmin_ref(1.0f, fmax_ref(-1.0f, ((1.0f + kRescale) * nv) + kRescale));
        W[t] = nv;
        a.E[t] = export_byte(nv);
        a.fl[p] = fabsf(nv - u4) > kTemporalThreshold ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// k_mark
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mark(KArgs a, int q)
{
    __shared__ uint32_t s_red[kBlock / 64];
    const uint32_t n = a.ctl->n[q];
    const bool few = path_is_few(n, a.g.N);
    const uint32_t* __restrict__ V = a.list[q];
    const float* __restrict__ Wn = a.phi[q ^ 1];  // freshly written buffer
    const bool ref = a.prm->refMode != 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    for (uint32_t p = gid; p < n; p += stride) {
        const uint32_t t = V[p];
        uint32_t m = 0;
        uint32_t nb[7], inb = 0;
        if ((!ref || (p & 3u) != 3u) && a.fl[p]) {
            neighbors(a.g, t, nb, &inb);
            const float u4 = Wn[t];
#pragma unroll
            for (int e = 0; e < 7; ++e) {
                if (e == 3) continue;
                if ((inb >> e) & 1u) {
                    const float un = Wn[nb[e]];
                    if (!(fabsf(u4 - un) <= kTemporalThreshold)) m |= 1u << e;
                }
            }
            if (m) m |= 8u;
        }
        a.mlist[p] = (uint8_t)m;
        if (m) {
            c += __popc(m);
            if (few) {
                a.M[t] = (uint8_t)m;
            } else {
#pragma unroll
                for (int e = 0; e < 7; ++e)
                    if ((m >> e) & 1u) a.F[nb[e]] = 1;
            }
        }
    }
    // reset the look-back state for k_compact
    const uint64_t slots = few ? 7ull * n : a.g.N;
    const uint32_t ntiles = (uint32_t)((slots + kScanTile - 1) / kScanTile);
    for (uint32_t i = gid; i < ntiles; i += stride) a.tstate[i] = 0;
    if (gid == 0) a.ctl->ticket = 0;
    // candidate count
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += s_red[w];
        if (tot) atomicAdd((unsigned long long*)&a.ctl->c_acc, (unsigned long long)tot);
    }
}

// ---------------------------------------------------------------------------
// k_compact
// ---------------------------------------------------------------------------
// init_mode: dense compaction of F into list[q^1] (initialisation), no stats.
__global__ __launch_bounds__(kBlock) void k_compact(KArgs a, int q, int init_mode)
{
    __shared__ uint32_t s_stage[kScanTile];
    __shared__ uint32_t s_cnt[kItems * 4];
    __shared__ uint32_t s_excl[kItems * 4];
    __shared__ uint32_t s_tile, s_prefix, s_total;

    const uint32_t n = a.ctl->n[q];
    const bool dense = init_mode || !path_is_few(n, a.g.N);
    const uint64_t slots = dense ? a.g.N : 7ull * n;
    const uint32_t ntiles = (uint32_t)((slots + kScanTile - 1) / kScanTile);
    const uint32_t* __restrict__ V = a.list[q];
    uint32_t* __restrict__ out = a.list[q ^ 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt(lane);

    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&a.ctl->ticket, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles) break;

        uint64_t bal[kItems];
        uint32_t val[kItems];
        const uint64_t base = (uint64_t)tile * kScanTile;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint64_t s = base + (uint64_t)j * kBlock + threadIdx.x;
            bool keep = false;
            uint32_t v = 0;
            if (s < slots) {
                if (dense) {
                    if (a.F[s]) {
                        keep = true;
                        v = (uint32_t)s;
                        a.F[s] = 0;
                    }
                } else {
                    // slot s = e * n + p (list-major order of the concatenation)
                    const uint32_t e = (uint32_t)(s >= n) + (uint32_t)(s >= 2ull * n) +
                                       (uint32_t)(s >= 3ull * n) + (uint32_t)(s >= 4ull * n) +
                                       (uint32_t)(s >= 5ull * n) + (uint32_t)(s >= 6ull * n);
                    const uint32_t p = (uint32_t)(s - (uint64_t)e * n);
                    if ((a.mlist[p] >> e) & 1u) {
                        const uint32_t t = V[p];
                        uint32_t nb[7], inb;
                        neighbors(a.g, t, nb, &inb);
                        const uint32_t u = nb[e];
                        keep = true;
                        if (e > 0) {
                            // generators of u in earlier lists: list e' emits
                            // w' + d(e') = u, i.e. w' = u - d(e') = nb(u, 6 - e')
                            uint32_t un[7], uin;
                            neighbors(a.g, u, un, &uin);
                            for (uint32_t e2 = 0; e2 < e; ++e2) {
                                const uint32_t o = 6u - e2;
                                if (o != 3u && !((uin >> o) & 1u)) continue;
                                if ((a.M[un[o]] >> e2) & 1u) {
                                    keep = false;
                                    break;
                                }
                            }
                        }
                        v = u;
                    }
                }
            }
            bal[j] = __ballot(keep);
            val[j] = v;
            if (lane == 0) s_cnt[j * 4 + wave] = (uint32_t)__popcll(bal[j]);
        }
        __syncthreads();
        if (wave == 0) {
            const uint32_t c = s_cnt[lane];
            const uint32_t incl = wave_inclusive_scan(c, lane);
            s_excl[lane] = incl - c;
            const uint32_t total = __shfl(incl, 63, 64);
            const uint32_t prefix = lookback(a.tstate, tile, total, &a.ctl->error);
            if (lane == 0) {
                s_prefix = prefix;
                s_total = total;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if ((bal[j] >> lane) & 1ull) {
                const uint32_t pos = s_excl[j * 4 + wave] + (uint32_t)__popcll(bal[j] & lt);
                s_stage[pos] = val[j];
            }
        }
        __syncthreads();
        const uint32_t total = s_total, prefix = s_prefix;
        for (uint32_t i = threadIdx.x; i < total; i += kBlock) out[prefix + i] = s_stage[i];
        if (threadIdx.x == 0 && tile == ntiles - 1) {
            const uint32_t m = prefix + total;
            a.ctl->n[q ^ 1] = m;
            if (init_mode) {
                a.ctl->n[q] = 0;
            } else {
                const uint32_t k = a.ctl->iter;
                if (k < a.ctl->stats_cap) {
                    a.n_hist[k] = n;
                    a.c_hist[k] = a.ctl->c_acc;
                    a.m_hist[k] = m;
                }
                if (dense) a.ctl->many_iters += 1; else a.ctl->few_iters += 1;
                a.ctl->c_acc = 0;
                a.ctl->iter = k + 1;
            }
        }
        __syncthreads();
    }
    if (ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctl->n[q ^ 1] = 0;
        if (init_mode) {
            a.ctl->n[q] = 0;
        } else {
            const uint32_t k = a.ctl->iter;
            if (k < a.ctl->stats_cap) {
                a.n_hist[k] = n;
                a.c_hist[k] = a.ctl->c_acc;
                a.m_hist[k] = 0;
            }
            if (dense) a.ctl->many_iters += 1; else a.ctl->few_iters += 1;
            a.ctl->c_acc = 0;
            a.ctl->iter = k + 1;
        }
    }
}

// ---------------------------------------------------------------------------
// dense helpers (initialisation, readback)
// ---------------------------------------------------------------------------
__global__ void k_fill_f32(float* p, uint64_t count, float v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, uint64_t count, uint8_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// pad + tile the linear source (LoadVolumeHostPadded, Segmenter.cpp:1131-1204)
__global__ void k_pad_tile_source(Geom g, const uint8_t* lin, uint8_t* dst, uint32_t nx, uint32_t ny,
                                  uint32_t nz, uint32_t bytes, uint8_t pad)
{
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        decode_xyz(g, (uint32_t)t, &x, &y, &z);
        if (x < nx && y < ny && z < nz) {
            const uint64_t li = ((uint64_t)z * ny + y) * nx + x;
            for (uint32_t b = 0; b < bytes; ++b) dst[t * bytes + b] = lin[li * bytes + b];
        } else {
            for (uint32_t b = 0; b < bytes; ++b) dst[t * bytes + b] = pad;
        }
    }
}

__global__ void k_set_phi(KArgs a, const float* lin, uint32_t nx, uint32_t ny, uint32_t nz, float band)
{
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        decode_xyz(a.g, (uint32_t)t, &x, &y, &z);
        float v = 1.0f;
        if (x < nx && y < ny && z < nz) v = lin[((uint64_t)z * ny + y) * nx + x];
        a.phi[0][t] = v;
        a.phi[1][t] = v;
        a.E[t] = export_byte(v);
        a.F[t] = fabsf(v) < band ? 1 : 0;
        a.M[t] = 0;
    }
}

// InitializeLevelSetKernel (InitializeLevelSetVolume.cu:55-98)
__global__ void k_seed_pass(KArgs a, int rd, uint32_t sx, uint32_t sy, uint32_t sz, int32_t radius,
                            float aniso)
{
    const float* R = a.phi[rd];
    float* W = a.phi[rd ^ 1];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        decode_xyz(a.g, (uint32_t)t, &x, &y, &z);
        const int32_t vx = (int32_t)(sx - x);
        const int32_t vy = (int32_t)(sy - y);
        const float vz = ((float)sz - (float)z) * aniso;
        const float d2 = (float)vx * (float)vx + (float)vy * (float)vy + vz * vz;
        const float d = __builtin_sqrtf(d2);
        float c = 1.0f;
        if (d < (float)radius) c = -1.0f;
        else if (d < (float)radius + kSmoothDistance) c = -1.0f + ((d - (float)radius) / (kSmoothDistance / 2.0f));
        float nv = R[t] + c - 1.0f;
        nv = nv < -1.0f ? -1.0f : (nv > 1.0f ? 1.0f : nv);
        W[t] = nv;
        a.E[t] = export_byte(nv);
    }
}

// InitializeActiveElementsVolumeKernel (InitializeActiveElementsVolume.cu:37-74)
__global__ void k_keep_flags(KArgs a, int rd)
{
    const float* R = a.phi[rd];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t nb[7], inb;
        neighbors(a.g, (uint32_t)t, nb, &inb);
        const float u4 = R[t];
        bool flat = true;
#pragma unroll
        for (int e = 0; e < 7; ++e) {
            if (e == 3) continue;
            if (!(fabsf(u4 - R[nb[e]]) <= kSpatialThreshold)) flat = false;
        }
        a.F[t] = flat ? 0 : 1;
        a.M[t] = 0;
    }
}

// AddLevelSetKernel (InitializeLevelSetVolume.cu:151-169)
__device__ inline int8_t add_export(int8_t ra, int8_t rb)
{
    const float av = (float)ra / 127.0f;
    const float bv = (float)rb / 127.0f;
    float r = av + bv - 1.0f;
    r = r < -1.0f ? -1.0f : (r > 1.0f ? 1.0f : r);
    return (int8_t)(int32_t)floorf(r * 127.0f);
}

__global__ void k_freeze(KArgs a)
{
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.g.N;
         t += (uint64_t)gridDim.x * blockDim.x)
        a.frozen[t] = add_export(a.frozen[t], a.E[t]);
}

__global__ void k_get_phi(KArgs a, int which, float* out, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint64_t cnt = (uint64_t)nx * ny * nz;
    const uint32_t X = a.g.X, Y = a.g.Y;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % nx);
        const uint32_t y = (uint32_t)((i / nx) % ny);
        const uint32_t z = (uint32_t)(i / ((uint64_t)nx * ny));
        const uint32_t t = (((z >> 4) * (Y >> 4) + (y >> 4)) * (X >> 4) + (x >> 4)) * 4096u +
                           ((z & 15u) * 16u + (y & 15u)) * 16u + (x & 15u);
        out[i] = a.phi[which][t];
    }
}

// GetSaveableSegmentation (Segmenter.cpp:1216-1286): export unioned with the
// frozen export, predicate <= 0, unpadded (UnpadVolume, :1288-1337)
__global__ void k_get_mask(KArgs a, uint8_t* out, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint64_t cnt = (uint64_t)nx * ny * nz;
    const uint32_t X = a.g.X, Y = a.g.Y;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % nx);
        const uint32_t y = (uint32_t)((i / nx) % ny);
        const uint32_t z = (uint32_t)(i / ((uint64_t)nx * ny));
        const uint32_t t = (((z >> 4) * (Y >> 4) + (y >> 4)) * (X >> 4) + (x >> 4)) * 4096u +
                           ((z & 15u) * 16u + (y & 15u)) * 16u + (x & 15u);
        out[i] = add_export(a.E[t], a.frozen[t]) <= 0 ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// host context
// ---------------------------------------------------------------------------
static thread_local std::string g_create_error;

struct Graphs {
    hipGraphExec_t chunk[2] = {nullptr, nullptr};  // kGraphIters iterations from parity 0 / 1
    int chunk_iters = 0;
};

}  // namespace lsseg

using namespace lsseg;

struct lsseg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t nx = 0, ny = 0, nz = 0;
    int32_t dtype = 0;
    uint32_t S = 1;
    float z_aniso = 1.0f;
    int q = 0;                        // current parity
    uint32_t grid = 1024;             // blocks for grid-stride kernels
    KArgs a{};
    DevParams* d_prm = nullptr;
    DevParams h_prm{};
    bool prm_valid = false;
    DevCtl* h_ctl = nullptr;          // pinned mirror
    uint32_t stats_cap = 0;
    uint64_t last_iters = 0;
    Graphs graphs;
    struct Checkpoint {
        float* phi[2] = {nullptr, nullptr};
        int8_t* E = nullptr;
        uint32_t* list = nullptr;
        uint32_t n = 0;
        bool valid = false;
    } ckpt[LSSEG_MAX_CHECKPOINTS];
    bool profile = false;               // record per-kernel events
    std::vector<hipEvent_t> prof_events;
    size_t prof_next = 0;
    std::vector<void*> allocs;
    std::string err;
};

namespace {

lsseg_status fail(lsseg_ctx* c, lsseg_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_error = buf;
    return st;
}

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((ctx), LSSEG_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

lsseg_status dalloc(lsseg_ctx* c, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, LSSEG_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs.push_back(*p);
    return LSSEG_OK;
}

void dfree(lsseg_ctx* c, void* p)
{
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it != c->allocs.end()) c->allocs.erase(it);
    (void)hipFree(p);
}

void destroy_graphs(lsseg_ctx* c)
{
    for (auto& g : c->graphs.chunk) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    c->graphs.chunk_iters = 0;
}

uint32_t dense_grid(uint64_t count)
{
    const uint64_t b = (count + kBlock - 1) / kBlock;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 65535));
}

int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

void launch_update(lsseg_ctx* c, int q)
{
    dim3 grid(c->grid), block(kBlock);
    switch (c->dtype) {
    case LSSEG_U8: hipLaunchKernelGGL((k_update<uint8_t, false>), grid, block, 0, c->stream, c->a, q); break;
    case LSSEG_I8: hipLaunchKernelGGL((k_update<int8_t, false>), grid, block, 0, c->stream, c->a, q); break;
    case LSSEG_U16: hipLaunchKernelGGL((k_update<uint16_t, false>), grid, block, 0, c->stream, c->a, q); break;
    case LSSEG_I16: hipLaunchKernelGGL((k_update<int16_t, false>), grid, block, 0, c->stream, c->a, q); break;
    case LSSEG_U32: hipLaunchKernelGGL((k_update<uint32_t, false>), grid, block, 0, c->stream, c->a, q); break;
    case LSSEG_I32: hipLaunchKernelGGL((k_update<int32_t, false>), grid, block, 0, c->stream, c->a, q); break;
    default: hipLaunchKernelGGL((k_update<float, true>), grid, block, 0, c->stream, c->a, q); break;
    }
}

void prof_mark(lsseg_ctx* c)
{
    if (c->profile && c->prof_next < c->prof_events.size())
        (void)hipEventRecord(c->prof_events[c->prof_next++], c->stream);
}

void launch_iteration(lsseg_ctx* c, int q)
{
    prof_mark(c);
    launch_update(c, q);
    prof_mark(c);
    hipLaunchKernelGGL(k_mark, dim3(c->grid), dim3(kBlock), 0, c->stream, c->a, q);
    prof_mark(c);
    hipLaunchKernelGGL(k_compact, dim3(c->grid), dim3(kBlock), 0, c->stream, c->a, q, 0);
    prof_mark(c);
}

// dense compaction of F into list[target]; list[target^1] count := 0
lsseg_status dense_init_list(lsseg_ctx* c, int target)
{
    const uint64_t ntiles = (c->a.g.N + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_fill_u64, dim3(dense_grid(ntiles)), dim3(kBlock), 0, c->stream, c->a.tstate,
                       ntiles, (uint64_t)0);
    HIP_TRY(c, hipMemsetAsync(&c->a.ctl->ticket, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_compact, dim3(c->grid), dim3(kBlock), 0, c->stream, c->a, target ^ 1, 1);
    HIP_TRY(c, hipGetLastError());
    c->q = target;
    return LSSEG_OK;
}

lsseg_status build_graphs(lsseg_ctx* c, int iters)
{
    destroy_graphs(c);
    for (int q0 = 0; q0 < 2; ++q0) {
        hipGraph_t g = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int q = q0;
        for (int i = 0; i < iters; ++i) {
            launch_iteration(c, q);
            q ^= 1;
        }
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&c->graphs.chunk[q0], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    c->graphs.chunk_iters = iters;
    return LSSEG_OK;
}

lsseg_status check_ctx(const lsseg_ctx* c)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    return LSSEG_OK;
}

lsseg_status sync_and_check(lsseg_ctx* c)
{
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->h_ctl->error)
        return fail(c, LSSEG_ERR_DEVICE, "device error flag 0x%x (look-back wait exceeded its bound)",
                    c->h_ctl->error);
    return LSSEG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* lsseg_last_error(const lsseg_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

lsseg_status lsseg_create(const lsseg_volume_desc* desc, const void* host_volume, lsseg_ctx** out_ctx)
{
    if (!desc || !host_volume || !out_ctx) return fail(nullptr, LSSEG_ERR_INVALID_ARG, "null argument");
    *out_ctx = nullptr;
    if (desc->nx == 0 || desc->ny == 0 || desc->nz == 0)
        return fail(nullptr, LSSEG_ERR_INVALID_ARG, "zero dimension");
    if (desc->dtype < LSSEG_U8 || desc->dtype > LSSEG_F32)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "unsupported dtype %d", desc->dtype);
    const uint64_t X = (desc->nx + 15ull) / 16 * 16, Y = (desc->ny + 15ull) / 16 * 16,
                   Z = (desc->nz + 15ull) / 16 * 16;
    const uint64_t N = X * Y * Z;
    if (N >= (1ull << 32))
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "padded volume of %llu voxels exceeds the 32-bit index",
                    (unsigned long long)N);

    lsseg_ctx* c = new (std::nothrow) lsseg_ctx();
    if (!c) return fail(nullptr, LSSEG_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto bail = [&](lsseg_status st) {
        g_create_error = c->err;
        lsseg_destroy(c);
        return st;
    };
    c->nx = desc->nx;
    c->ny = desc->ny;
    c->nz = desc->nz;
    c->dtype = desc->dtype;
    c->S = (desc->dtype <= LSSEG_I8) ? 1 : (desc->dtype <= LSSEG_I16 ? 2 : 4);
    c->z_aniso = desc->z_aniso;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        c->err = std::string("no HIP device: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    c->grid = (uint32_t)env_int("LSSEG_GRID_BLOCKS", std::max(1, cus) * 4);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        c->err = std::string("stream/event creation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    if (hipHostMalloc((void**)&c->h_ctl, sizeof(DevCtl), hipHostMallocDefault) != hipSuccess) {
        c->err = "pinned allocation failed";
        return bail(LSSEG_ERR_OUT_OF_MEMORY);
    }
    memset(c->h_ctl, 0, sizeof(DevCtl));

    Geom& g = c->a.g;
    g.X = (uint32_t)X;
    g.Y = (uint32_t)Y;
    g.Z = (uint32_t)Z;
    g.TX = g.X / 16;
    g.TY = g.Y / 16;
    g.TZ = g.Z / 16;
    g.sTY = g.TX * 4096u;
    g.sTZ = g.TX * g.TY * 4096u;
    g.N = N;
    g.dTX = make_fastdiv(g.TX);
    g.dTXY = make_fastdiv(g.TX * g.TY);

    void* p = nullptr;
    const uint64_t ntiles = N / kScanTile + 2;
    lsseg_status st;
#define ALLOC(field, type, count)                                      \
    if ((st = dalloc(c, &p, (size_t)(count) * sizeof(type))) != LSSEG_OK) \
        return bail(st);                                               \
    c->a.field = (type*)p;
    ALLOC(phi[0], float, N);
    ALLOC(phi[1], float, N);
    ALLOC(E, int8_t, N);
    ALLOC(frozen, int8_t, N);
    ALLOC(M, uint8_t, N);
    ALLOC(F, uint8_t, N);
    ALLOC(list[0], uint32_t, N);
    ALLOC(list[1], uint32_t, N);
    ALLOC(fl, uint8_t, N);
    ALLOC(mlist, uint8_t, N);
    ALLOC(tstate, uint64_t, ntiles);
    ALLOC(ctl, DevCtl, 1);
#undef ALLOC
    if ((st = dalloc(c, &p, (size_t)N * c->S)) != LSSEG_OK) return bail(st);
    c->a.src = p;
    if ((st = dalloc(c, &p, sizeof(DevParams))) != LSSEG_OK) return bail(st);
    c->d_prm = (DevParams*)p;
    c->a.prm = c->d_prm;

    // upload, pad and tile the source
    const uint64_t lin_bytes = (uint64_t)c->nx * c->ny * c->nz * c->S;
    void* lin = nullptr;
    if (hipMalloc(&lin, lin_bytes) != hipSuccess) {
        (void)hipGetLastError();
        c->err = "hipMalloc of the staging volume failed";
        return bail(LSSEG_ERR_OUT_OF_MEMORY);
    }
    const bool is_signed = c->dtype == LSSEG_I8 || c->dtype == LSSEG_I16 || c->dtype == LSSEG_I32;
    e = hipMemcpyAsync(lin, host_volume, lin_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_pad_tile_source, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, g,
                           (const uint8_t*)lin, (uint8_t*)c->a.src, c->nx, c->ny, c->nz, c->S,
                           (uint8_t)(is_signed ? 0x81 : 0x00));
        e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(lin);
    if (e != hipSuccess) {
        c->err = std::string("source upload failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    // Segmenter.cpp:150-154, 265-266: phi := 1, export := 127, scratch := 0
    const uint32_t dg = dense_grid(N);
    hipLaunchKernelGGL(k_fill_f32, dim3(dg), dim3(kBlock), 0, c->stream, c->a.phi[0], N, 1.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(dg), dim3(kBlock), 0, c->stream, c->a.phi[1], N, 1.0f);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.E, N, (uint8_t)127);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.frozen, N, (uint8_t)127);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.M, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.F, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.mlist, N, (uint8_t)0);
    e = hipMemsetAsync(c->a.ctl, 0, sizeof(DevCtl), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->err = std::string("initialisation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    c->q = 0;
    *out_ctx = c;
    return LSSEG_OK;
}

lsseg_status lsseg_destroy(lsseg_ctx* ctx)
{
    if (!ctx) return LSSEG_ERR_INVALID_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    destroy_graphs(ctx);
    for (hipEvent_t e : ctx->prof_events) (void)hipEventDestroy(e);
    ctx->prof_events.clear();
    for (void* p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LSSEG_OK;
}

lsseg_status lsseg_get_dims(const lsseg_ctx* ctx, uint32_t* X, uint32_t* Y, uint32_t* Z)
{
    if (!ctx || !X || !Y || !Z) return LSSEG_ERR_INVALID_ARG;
    *X = ctx->a.g.X;
    *Y = ctx->a.g.Y;
    *Z = ctx->a.g.Z;
    return LSSEG_OK;
}

lsseg_status lsseg_set_phi(lsseg_ctx* c, const float* phi0, float band)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (!phi0) return fail(c, LSSEG_ERR_INVALID_ARG, "phi0 is null");
    if (!(band > 0.0f)) return fail(c, LSSEG_ERR_INVALID_ARG, "band must be positive");
    const uint64_t cnt = (uint64_t)c->nx * c->ny * c->nz;
    float* lin = nullptr;
    if (hipMalloc((void**)&lin, cnt * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, LSSEG_ERR_OUT_OF_MEMORY, "staging allocation failed");
    }
    hipError_t e = hipMemcpyAsync(lin, phi0, cnt * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_set_phi, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a, lin,
                           c->nx, c->ny, c->nz, band);
        e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(lin);
    if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "set_phi failed: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a.mlist,
                       c->a.g.N, (uint8_t)0);
    lsseg_status st = dense_init_list(c, 0);
    if (st != LSSEG_OK) return st;
    return sync_and_check(c);
}

lsseg_status lsseg_seed_sphere(lsseg_ctx* c, float cx, float cy, float cz, uint32_t radius)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (!(cx >= 0.0f) || !(cy >= 0.0f) || !(cz >= 0.0f) || radius > (1u << 20))
        return fail(c, LSSEG_ERR_INVALID_ARG, "seed centre must be non-negative and the radius sane");
    const uint32_t sx = (uint32_t)cx, sy = (uint32_t)cy, sz = (uint32_t)cz;  // Segmenter.cpp:502-505
    const uint32_t dg = dense_grid(c->a.g.N);
    int rd = c->q;
    for (int pass = 0; pass < 2; ++pass) {  // Segmenter.cpp:525-531
        hipLaunchKernelGGL(k_seed_pass, dim3(dg), dim3(kBlock), 0, c->stream, c->a, rd, sx, sy, sz,
                           (int32_t)radius, c->z_aniso);
        rd ^= 1;
    }
    hipLaunchKernelGGL(k_keep_flags, dim3(dg), dim3(kBlock), 0, c->stream, c->a, rd);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.mlist, c->a.g.N, (uint8_t)0);
    HIP_TRY(c, hipGetLastError());
    // parity is unchanged after two swaps: the list goes to the read parity
    lsseg_status st = dense_init_list(c, rd);
    if (st != LSSEG_OK) return st;
    return sync_and_check(c);
}

}  // extern "C"

namespace {

// Validates the parameters and derives the device constants.
lsseg_status make_params(lsseg_ctx* c, const lsseg_params* prm, DevParams* out)
{
    if (!prm) return fail(c, LSSEG_ERR_INVALID_ARG, "params is null");
    if (prm->mode != LSSEG_MODE_REF && prm->mode != LSSEG_MODE_CLEAN)
        return fail(c, LSSEG_ERR_INVALID_ARG, "unknown mode %d", prm->mode);
    if (!std::isfinite(prm->T) || !std::isfinite(prm->eps) || !std::isfinite(prm->alpha) ||
        !std::isfinite(prm->dt))
        return fail(c, LSSEG_ERR_INVALID_ARG, "non-finite parameter");
    DevParams P{};
    if (c->dtype == LSSEG_F32) {
        if (!(prm->eps > 0.0f)) return fail(c, LSSEG_ERR_INVALID_ARG, "eps must be positive");
        P.negInv = -1.0f / prm->eps;
        P.Tf = prm->T;
    } else {
        // Segmenter.cpp:817-818: static_cast<int> of T and eps
        if (!(prm->eps >= 1.0f) || prm->eps > 2147483647.0f || std::fabs(prm->T) > 2147483647.0f)
            return fail(c, LSSEG_ERR_INVALID_ARG, "integer sources need 1 <= eps < 2^31 and |T| < 2^31");
        P.Ti = (int32_t)prm->T;
        const uint32_t eu = (uint32_t)(int32_t)prm->eps;
        P.negInv = -1.0f / (float)eu;  // ComputeDensityTerm, UpdateLevelSetVolume.cu:339
    }
    P.alpha = prm->alpha;
    P.oneMinusAlpha = 1.0f - prm->alpha;
    P.dtAbs = prm->dt < 0.0f ? -prm->dt : prm->dt;
    P.flipDensity = prm->dt < 0.0f ? 1 : 0;
    P.alphaPositive = prm->alpha > 0.0f ? 1 : 0;
    P.refMode = prm->mode == LSSEG_MODE_REF ? 1 : 0;
    *out = P;
    return LSSEG_OK;
}

lsseg_status upload_params(lsseg_ctx* c, const DevParams& P)
{
    c->h_prm = P;
    HIP_TRY(c, hipMemcpyAsync(c->d_prm, &c->h_prm, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
    return LSSEG_OK;
}

}  // namespace

extern "C" {

lsseg_status lsseg_iterate_reinit(lsseg_ctx* c, const lsseg_params* prm)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    DevParams P;
    lsseg_status st = make_params(c, prm, &P);
    if (st != LSSEG_OK) return st;
    if ((st = upload_params(c, P)) != LSSEG_OK) return st;
    launch_update(c, c->q);
    const int rd = c->q ^ 1;
    hipLaunchKernelGGL(k_keep_flags, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a, rd);
    HIP_TRY(c, hipGetLastError());
    if ((st = dense_init_list(c, rd)) != LSSEG_OK) return st;
    return sync_and_check(c);
}

lsseg_status lsseg_iterate(lsseg_ctx* c, const lsseg_params* prm, int32_t iters, lsseg_stats* stats)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (iters < 0) return fail(c, LSSEG_ERR_INVALID_ARG, "negative iteration count");
    DevParams P;
    lsseg_status pst = make_params(c, prm, &P);
    if (pst != LSSEG_OK) return pst;

    // per-iteration statistics capacity
    if ((uint64_t)iters > c->stats_cap) {
        const uint32_t cap = (uint32_t)std::max<int64_t>(iters, 1024);
        void *pn = nullptr, *pc = nullptr, *pm = nullptr;
        lsseg_status st;
        if ((st = dalloc(c, &pn, cap * sizeof(uint32_t))) != LSSEG_OK) return st;
        if ((st = dalloc(c, &pc, cap * sizeof(uint64_t))) != LSSEG_OK) return st;
        if ((st = dalloc(c, &pm, cap * sizeof(uint32_t))) != LSSEG_OK) return st;
        dfree(c, c->a.n_hist);
        dfree(c, c->a.c_hist);
        dfree(c, c->a.m_hist);
        c->a.n_hist = (uint32_t*)pn;
        c->a.c_hist = (uint64_t*)pc;
        c->a.m_hist = (uint32_t*)pm;
        c->stats_cap = cap;
        destroy_graphs(c);  // captured kernel arguments changed
    }
    if ((pst = upload_params(c, P)) != LSSEG_OK) return pst;
    // reset per-call counters (keep n[], ticket)
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->h_ctl->iter = 0;
    c->h_ctl->stats_cap = c->stats_cap;
    c->h_ctl->error = 0;
    c->h_ctl->c_acc = 0;
    c->h_ctl->few_iters = 0;
    c->h_ctl->many_iters = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl, c->h_ctl, sizeof(DevCtl), hipMemcpyHostToDevice, c->stream));

    const int chunk = env_int("LSSEG_GRAPH_ITERS", 32);
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    int remaining = iters;
    if (!c->profile && chunk > 1 && remaining >= chunk) {
        if (c->graphs.chunk_iters != chunk) {
            lsseg_status st = build_graphs(c, chunk);
            if (st != LSSEG_OK) return st;
        }
        while (remaining >= chunk) {
            HIP_TRY(c, hipGraphLaunch(c->graphs.chunk[c->q], c->stream));
            if (chunk & 1) c->q ^= 1;
            remaining -= chunk;
        }
    }
    for (; remaining > 0; --remaining) {
        launch_iteration(c, c->q);
        c->q ^= 1;
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    lsseg_status st = sync_and_check(c);
    if (st != LSSEG_OK) return st;
    c->last_iters = (uint64_t)iters;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iterations = (uint64_t)iters;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        stats->device_ms = ms;
        stats->active_final = c->h_ctl->n[c->q];
        stats->few_path_iters = c->h_ctl->few_iters;
        stats->many_path_iters = c->h_ctl->many_iters;
        if (iters > 0) {
            std::vector<uint32_t> hn(iters), hm(iters);
            std::vector<uint64_t> hc(iters);
            HIP_TRY(c, hipMemcpy(hn.data(), c->a.n_hist, iters * sizeof(uint32_t), hipMemcpyDeviceToHost));
            HIP_TRY(c, hipMemcpy(hc.data(), c->a.c_hist, iters * sizeof(uint64_t), hipMemcpyDeviceToHost));
            HIP_TRY(c, hipMemcpy(hm.data(), c->a.m_hist, iters * sizeof(uint32_t), hipMemcpyDeviceToHost));
            for (int i = 0; i < iters; ++i) {
                stats->sum_active += hn[i];
                stats->sum_candidates += hc[i];
                stats->sum_new_active += hm[i];
            }
        }
    }
    return LSSEG_OK;
}

lsseg_status lsseg_get_iteration_counts(const lsseg_ctx* c, uint32_t* n, uint64_t* cc, uint32_t* m,
                                        uint64_t cap, uint64_t* count)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    const uint64_t k = std::min<uint64_t>(cap, c->last_iters);
    if (count) *count = c->last_iters;
    if (k) {
        if (n) HIP_TRY(cm, hipMemcpy(n, c->a.n_hist, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (cc) HIP_TRY(cm, hipMemcpy(cc, c->a.c_hist, k * sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (m) HIP_TRY(cm, hipMemcpy(m, c->a.m_hist, k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    if (!n && !cc && !m) return LSSEG_OK;  // count query
    return cap < c->last_iters ? fail(cm, LSSEG_ERR_CAPACITY, "%llu iterations, capacity %llu",
                                      (unsigned long long)c->last_iters, (unsigned long long)cap)
                               : LSSEG_OK;
}

lsseg_status lsseg_get_active_count(const lsseg_ctx* c, uint64_t* n)
{
    if (!c || !n) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    uint32_t v = 0;
    HIP_TRY(cm, hipMemcpyAsync(&v, &c->a.ctl->n[c->q], sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(cm, hipStreamSynchronize(c->stream));
    *n = v;
    return LSSEG_OK;
}

lsseg_status lsseg_get_active(const lsseg_ctx* c, uint32_t* out, uint64_t cap, uint64_t* n)
{
    if (!c || !n || (!out && cap)) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    uint64_t cnt = 0;
    lsseg_status st = lsseg_get_active_count(c, &cnt);
    if (st != LSSEG_OK) return st;
    *n = cnt;
    const uint64_t k = std::min(cap, cnt);
    if (k) HIP_TRY(cm, hipMemcpy(out, c->a.list[c->q], k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return cap < cnt ? fail(cm, LSSEG_ERR_CAPACITY, "list has %llu entries, capacity %llu",
                            (unsigned long long)cnt, (unsigned long long)cap)
                     : LSSEG_OK;
}

lsseg_status lsseg_get_phi(const lsseg_ctx* c, int32_t which, float* out)
{
    if (!c || !out || (which != 0 && which != 1)) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    const uint64_t cnt = (uint64_t)c->nx * c->ny * c->nz;
    float* d = nullptr;
    if (hipMalloc((void**)&d, cnt * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(cm, LSSEG_ERR_OUT_OF_MEMORY, "staging allocation failed");
    }
    hipLaunchKernelGGL(k_get_phi, dim3(dense_grid(cnt)), dim3(kBlock), 0, c->stream, c->a,
                       which == 0 ? c->q : c->q ^ 1, d, c->nx, c->ny, c->nz);
    hipError_t e = hipMemcpyAsync(out, d, cnt * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(cm, LSSEG_ERR_DEVICE, "get_phi failed: %s", hipGetErrorString(e));
    return LSSEG_OK;
}

lsseg_status lsseg_get_mask(const lsseg_ctx* c, uint8_t* out)
{
    if (!c || !out) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    const uint64_t cnt = (uint64_t)c->nx * c->ny * c->nz;
    uint8_t* d = nullptr;
    if (hipMalloc((void**)&d, cnt) != hipSuccess) {
        (void)hipGetLastError();
        return fail(cm, LSSEG_ERR_OUT_OF_MEMORY, "staging allocation failed");
    }
    hipLaunchKernelGGL(k_get_mask, dim3(dense_grid(cnt)), dim3(kBlock), 0, c->stream, c->a, d, c->nx,
                       c->ny, c->nz);
    hipError_t e = hipMemcpyAsync(out, d, cnt, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(cm, LSSEG_ERR_DEVICE, "get_mask failed: %s", hipGetErrorString(e));
    return LSSEG_OK;
}

lsseg_status lsseg_freeze(lsseg_ctx* c)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_freeze, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LSSEG_OK;
}

lsseg_status lsseg_clear(lsseg_ctx* c, int32_t all)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    const uint64_t N = c->a.g.N;
    const uint32_t dg = dense_grid(N);
    hipLaunchKernelGGL(k_fill_f32, dim3(dg), dim3(kBlock), 0, c->stream, c->a.phi[0], N, 1.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(dg), dim3(kBlock), 0, c->stream, c->a.phi[1], N, 1.0f);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.E, N, (uint8_t)127);
    if (all)
        hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.frozen, N,
                           (uint8_t)127);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.M, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.F, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.mlist, N, (uint8_t)0);
    HIP_TRY(c, hipMemsetAsync(c->a.ctl->n, 0, 2 * sizeof(uint32_t), c->stream));
    HIP_TRY(c, hipGetLastError());
    return sync_and_check(c);
}

lsseg_status lsseg_get_state(const lsseg_ctx* c, float* phi_read, float* phi_write, int8_t* E,
                             uint32_t* list, uint64_t list_cap, uint64_t* n)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    const uint64_t N = c->a.g.N;
    if (phi_read) HIP_TRY(cm, hipMemcpy(phi_read, c->a.phi[c->q], N * sizeof(float), hipMemcpyDeviceToHost));
    if (phi_write)
        HIP_TRY(cm, hipMemcpy(phi_write, c->a.phi[c->q ^ 1], N * sizeof(float), hipMemcpyDeviceToHost));
    if (E) HIP_TRY(cm, hipMemcpy(E, c->a.E, N, hipMemcpyDeviceToHost));
    if (n || list) {
        uint64_t cnt = 0;
        if (list && !n) return fail(cm, LSSEG_ERR_INVALID_ARG, "list requires n");
        return lsseg_get_active(c, list, list ? list_cap : 0, n ? n : &cnt);
    }
    return LSSEG_OK;
}

lsseg_status lsseg_set_state(lsseg_ctx* c, const float* phi_read, const float* phi_write, const int8_t* E,
                             const uint32_t* list, uint64_t n)
{
    if (!c || !phi_read || !phi_write || !E || (!list && n)) return LSSEG_ERR_INVALID_ARG;
    const uint64_t N = c->a.g.N;
    if (n > N) return fail(c, LSSEG_ERR_CAPACITY, "list of %llu entries exceeds the volume", (unsigned long long)n);
    std::vector<uint8_t> seen(N, 0);
    for (uint64_t i = 0; i < n; ++i) {
        if (list[i] >= N) return fail(c, LSSEG_ERR_INVALID_ARG, "list entry %llu out of range", (unsigned long long)i);
        if (seen[list[i]]) return fail(c, LSSEG_ERR_INVALID_ARG, "list entry %llu is a duplicate", (unsigned long long)i);
        seen[list[i]] = 1;
    }
    c->q = 0;
    HIP_TRY(c, hipMemcpy(c->a.phi[0], phi_read, N * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->a.phi[1], phi_write, N * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->a.E, E, N, hipMemcpyHostToDevice));
    if (n) HIP_TRY(c, hipMemcpy(c->a.list[0], list, n * sizeof(uint32_t), hipMemcpyHostToDevice));
    const uint32_t dg = dense_grid(N);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.M, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.F, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.mlist, N, (uint8_t)0);
    uint32_t counts[2] = {(uint32_t)n, 0u};
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl->n, counts, sizeof(counts), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipGetLastError());
    return sync_and_check(c);
}

lsseg_status lsseg_iterate_profiled(lsseg_ctx* c, const lsseg_params* prm, int32_t iters, lsseg_stats* stats,
                                    lsseg_kernel_times* times)
{
    if (!c || !times) return LSSEG_ERR_INVALID_ARG;
    if (iters < 0) return fail(c, LSSEG_ERR_INVALID_ARG, "negative iteration count");
    const size_t need = (size_t)iters * 4;
    while (c->prof_events.size() < need) {
        hipEvent_t e;
        HIP_TRY(c, hipEventCreate(&e));
        c->prof_events.push_back(e);
    }
    c->profile = true;
    c->prof_next = 0;
    lsseg_status st = lsseg_iterate(c, prm, iters, stats);
    c->profile = false;
    if (st != LSSEG_OK) return st;
    memset(times, 0, sizeof(*times));
    for (int32_t i = 0; i < iters; ++i) {
        float a = 0, b = 0, d = 0;
        HIP_TRY(c, hipEventElapsedTime(&a, c->prof_events[4 * i + 0], c->prof_events[4 * i + 1]));
        HIP_TRY(c, hipEventElapsedTime(&b, c->prof_events[4 * i + 1], c->prof_events[4 * i + 2]));
        HIP_TRY(c, hipEventElapsedTime(&d, c->prof_events[4 * i + 2], c->prof_events[4 * i + 3]));
        times->update_ms += a;
        times->mark_ms += b;
        times->compact_ms += d;
    }
    times->iterations = (uint64_t)iters;
    return LSSEG_OK;
}

lsseg_status lsseg_checkpoint_save(lsseg_ctx* c, int32_t slot)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (slot < 0 || slot >= LSSEG_MAX_CHECKPOINTS) return fail(c, LSSEG_ERR_INVALID_ARG, "bad checkpoint slot %d", slot);
    auto& k = c->ckpt[slot];
    const uint64_t N = c->a.g.N;
    if (!k.phi[0]) {
        void* p = nullptr;
        lsseg_status st;
        if ((st = dalloc(c, &p, N * sizeof(float))) != LSSEG_OK) return st;
        k.phi[0] = (float*)p;
        if ((st = dalloc(c, &p, N * sizeof(float))) != LSSEG_OK) return st;
        k.phi[1] = (float*)p;
        if ((st = dalloc(c, &p, N)) != LSSEG_OK) return st;
        k.E = (int8_t*)p;
        if ((st = dalloc(c, &p, N * sizeof(uint32_t))) != LSSEG_OK) return st;
        k.list = (uint32_t*)p;
    }
    uint64_t n = 0;
    lsseg_status st = lsseg_get_active_count(c, &n);
    if (st != LSSEG_OK) return st;
    HIP_TRY(c, hipMemcpyAsync(k.phi[0], c->a.phi[c->q], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(k.phi[1], c->a.phi[c->q ^ 1], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(k.E, c->a.E, N, hipMemcpyDeviceToDevice, c->stream));
    if (n) HIP_TRY(c, hipMemcpyAsync(k.list, c->a.list[c->q], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    k.n = (uint32_t)n;
    k.valid = true;
    return LSSEG_OK;
}

lsseg_status lsseg_checkpoint_restore(lsseg_ctx* c, int32_t slot)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (slot < 0 || slot >= LSSEG_MAX_CHECKPOINTS || !c->ckpt[slot].valid)
        return fail(c, LSSEG_ERR_STATE, "checkpoint slot %d is empty", slot);
    auto& k = c->ckpt[slot];
    const uint64_t N = c->a.g.N;
    c->q = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.phi[0], k.phi[0], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->a.phi[1], k.phi[1], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->a.E, k.E, N, hipMemcpyDeviceToDevice, c->stream));
    if (k.n) HIP_TRY(c, hipMemcpyAsync(c->a.list[0], k.list, k.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    // per-voxel emission masks must be zero outside an iteration in flight
    hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, c->a.M, N, (uint8_t)0);
    uint32_t* counts = &c->h_ctl->n[0];
    counts[0] = k.n;
    counts[1] = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl->n, counts, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LSSEG_OK;
}

lsseg_status lsseg_get_device_buffers(const lsseg_ctx* c, void** phi_read, void** phi_write, void** stream)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (phi_read) *phi_read = c->a.phi[c->q];
    if (phi_write) *phi_write = c->a.phi[c->q ^ 1];
    if (stream) *stream = (void*)c->stream;
    return LSSEG_OK;
}

}  // extern "C"
