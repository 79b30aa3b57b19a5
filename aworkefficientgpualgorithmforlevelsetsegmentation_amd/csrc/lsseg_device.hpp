// lsseg_device.hpp -- device code of the sparse level-set engine (gfx950).
//
// Reference paths in comments are relative to code/gpuseg/code/ of the
// reference repository.
//
// The active list V (the reference's order, which decides the p % 4 == 3 rule
// and the next list's order) is kept together with S, the same voxels in a
// spatially coherent order, and pos[u], the position of u in V.  The list order
// scatters neighbouring voxels over the whole volume after a few hundred
// iterations, so every per-voxel step runs over S, where consecutive lanes
// touch neighbouring voxels, and only the compaction walks V.  The steps of
// iteration k (all counts in device memory):
//   update   over S: the 7-point update of phi (UpdateLevelSetVolume.cu:222-342)
//            from the read buffer into the write buffer, plus the export byte;
//   mark     over S: the temporal AND spatial test on the new phi
//            (UpdateActiveElements.cu:75-103), positions p % 4 == 3 skipped
//            (UpdateActiveElementsVolume.cu:329), candidates recorded, the
//            emission mask kept per S position;
//   keep     over S: first-wins duplicate removal (FilterDuplicates.cu:42-66),
//            one byte of kept lists per list position, the kept voxels in S
//            order (the next S), the next count and the statistics;
//   compact  over the candidate slots in list order: stable compaction into the
//            next list (Segmenter.cpp:883-1019, Cuda.cpp:699-758) and its
//            positions.
// The compaction of iteration k touches nothing that the update of k + 1
// reads or writes, so both run in one launch (k_update_fused): three launches
// per iteration.  Keep and compaction are single-pass scans with one chunk per
// workgroup, after the decoupled look-back of Merrill & Garland ("Single-pass
// Parallel Prefix Scan with Decoupled Look-back", 2016) but with two levels
// (chunk aggregates and totals of groups of 64 chunks, wave_group_scan)
// instead of a chain of inclusive prefixes.
//
// Emission record C (8 bytes per voxel): byte e of C[u] holds the iteration's
// stamp iff u was emitted into candidate list e (0:-z 1:-y 2:-x 3:self 4:+x
// 5:+y 6:+z, UpdateActiveElementsVolume.cu:168-174).  Over the list-major
// concatenation of the seven lists, first-wins keeps candidate (e, p) iff the
// lowest byte of C[u] holding the stamp is e; the dense path keeps the kept
// voxels (marked in memb by the keep step) in tiled order.  The stamp (1..255,
// DevCtl::stamp) advances with every compaction, so the bytes of earlier
// iterations need no clearing: the host clears the whole record and restarts
// the stamp before it would pass 255.  The first update of a call still
// clears C at its voxels, which removes the flags an initialisation leaves.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsseg {

// Config.hpp:55-66
constexpr float kTemporalThreshold = 0.008f;  // TEMPORAL_DERIVATIVE_THRESHOLD
constexpr float kSpatialThreshold = 0.001f;   // SPATIAL_DERIVATIVE_THRESHOLD
constexpr float kRescale = 0.01f;             // LEVEL_SET_RESCALE_AMOUNT
constexpr float kSmoothDistance = 4.0f;       // LEVEL_SET_INITIALIZE_SMOOTH_DISTANCE
constexpr float kConnTol = 0.1f;              // UpdateLevelSetVolume.cu:282-287

constexpr int kBlock = 256;                   // threads per workgroup (4 waves of 64)
constexpr int kItems = 8;                     // compaction slots per thread
constexpr int kScanTile = kBlock * kItems;    // slots per look-back tile
constexpr int kKeepItems = 1;                 // S positions per thread in the keep step (measured:
                                              // 2 costs 4% on the phantom, 4 costs 10%)
constexpr int kKeepItemsBatch = 2;            // the default of the batched keep step (autotuned: 1
                                              // cost 5% on 64 volumes, whose scans share one grid)
constexpr int kKeepTile = kBlock * kKeepItems;  // the smallest keep tile (state sizing)
constexpr int kPosItems = 4;                  // list positions per thread in a few-path compaction (<= 4;
                                              // 2 measured 4% slower on the phantom, 9% on the vessels);
                                              // lists of at most Geom::small_n positions take one
constexpr int kPosTile = kBlock * kPosItems;
constexpr uint32_t kSpecTiles = 256;          // default of Geom::spec_tiles
constexpr uint32_t kSpinLimit = 1u << 24;     // bounded look-back wait

enum : uint32_t { kErrSpin = 1u, kErrOverflow = 2u, kErrCount = 8u, kErrCheck = 16u, kErrStamp = 32u };

enum KMode : int {
    kUpdate = 0,   // update of the current list
    kMark = 1,     // marking of the current list
    kCompact = 2,  // duplicate removal + compaction into the next list, statistics
    kInit = 3      // dense compaction of C only (initialisation), no statistics
};

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------
struct FastDiv {
    uint32_t d, m, s;
};

// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = 2^32 (2^s - d) / d + 1
inline FastDiv make_fastdiv(uint32_t d)
{
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
    uint32_t sTY, sTZ;       // tiled strides of one tile row (y) and one tile slab (z)
    uint64_t N;
    FastDiv dTX, dTXY;
    uint32_t small_n;        // few-path compactions of at most this many list positions take
                             // tiles of kBlock positions, the others of kPosTile (0: never)
    uint32_t spec_tiles;     // workgroups of a compaction that load their tile with the counts
    uint32_t pack_counts;    // 1: a few-path compaction chunk publishes its seven counts two to a
                             // look-back word where they fit below 2^16; 0: one count per word
};

// Voxel ids.  Voxel (x, y, z) of the padded volume is stored at
//   id = tile * 4096 + m,  tile = (tz * TY + ty) * TX + tx,
// m the Morton code of its coordinates (lx, ly, lz) inside the 16^3 tile: bit
// 3k of m is bit k of lx, bit 3k + 1 of ly, bit 3k + 2 of lz.  A 32-byte
// sector of a float array then holds a 2x2x2 block and a 128-byte line a
// 4x4x2 block, so the gathers around the active surface touch fewer sectors
// than with rows of 16 whatever the surface's orientation.  The reference's
// tiled index tau (the same tile, then lz * 256 + ly * 16 + lx) still decides
// the dense path's list order, and the C ABI speaks tau; tau_of_id and
// id_of_tau convert.
constexpr uint32_t kMx = 0x249u, kMy = 0x492u, kMz = 0x924u;  // the bits of lx, ly, lz in m

__host__ __device__ inline uint32_t spread4(uint32_t v)
{
    v = (v | (v << 4)) & 0x0C3u;
    return (v | (v << 2)) & 0x249u;
}

__host__ __device__ inline uint32_t compact4(uint32_t v)
{
    v &= 0x249u;
    v = (v | (v >> 2)) & 0x0C3u;
    return (v | (v >> 4)) & 0x00Fu;
}

__host__ __device__ inline uint32_t id_of_tau(uint32_t t)
{
    return (t & ~4095u) | spread4(t & 15u) | (spread4((t >> 4) & 15u) << 1) | (spread4((t >> 8) & 15u) << 2);
}

__host__ __device__ inline uint32_t tau_of_id(uint32_t v)
{
    const uint32_t m = v & 4095u;
    return (v & ~4095u) | (compact4(m >> 2) << 8) | (compact4(m >> 1) << 4) | compact4(m);
}

// one step along the axis whose bits in m are M (the other axes' bits stay; the
// zero gaps between M's bits carry the borrow or the carry)
__host__ __device__ inline uint32_t morton_dec(uint32_t m, uint32_t M) { return (((m & M) - 1u) & M) | (m & ~M); }
__host__ __device__ inline uint32_t morton_inc(uint32_t m, uint32_t M) { return (((m | ~M) + 1u) & M) | (m & ~M); }

// The six face neighbours in list order; an out-of-bounds neighbour is the
// voxel itself (OffsetCoordinatesUnsafe, Coordinates.cu:157-184) and bit e of
// *inb tells whether neighbour e is inside the volume.  Across a tile face the
// axis' in-tile coordinate goes from 0 to 15 (all its bits set) or back.
__host__ __device__ inline void neighbors(const Geom& g, uint32_t t, uint32_t nb[7], uint32_t* inb)
{
    const uint32_t m = t & 4095u, hi = t - m;
    const uint32_t xm = m & kMx, ym = m & kMy, zm = m & kMz;
    const uint32_t tile = t >> 12;
    const uint32_t qy = fdiv(tile, g.dTX);
    const uint32_t tx = tile - qy * g.TX;
    const uint32_t tz = fdiv(tile, g.dTXY);
    const uint32_t ty = qy - tz * g.TY;
    uint32_t in = 0;
    nb[0] = zm ? hi | morton_dec(m, kMz) : (tz ? t - g.sTZ + kMz : t);
    in |= (zm | tz) ? 1u : 0u;
    nb[1] = ym ? hi | morton_dec(m, kMy) : (ty ? t - g.sTY + kMy : t);
    in |= (ym | ty) ? 2u : 0u;
    nb[2] = xm ? hi | morton_dec(m, kMx) : (tx ? t - 4096u + kMx : t);
    in |= (xm | tx) ? 4u : 0u;
    nb[3] = t;
    nb[4] = xm != kMx ? hi | morton_inc(m, kMx) : (tx + 1u < g.TX ? t + 4096u - kMx : t);
    in |= (xm != kMx || tx + 1u < g.TX) ? 16u : 0u;
    nb[5] = ym != kMy ? hi | morton_inc(m, kMy) : (ty + 1u < g.TY ? t + g.sTY - kMy : t);
    in |= (ym != kMy || ty + 1u < g.TY) ? 32u : 0u;
    nb[6] = zm != kMz ? hi | morton_inc(m, kMz) : (tz + 1u < g.TZ ? t + g.sTZ - kMz : t);
    in |= (zm != kMz || tz + 1u < g.TZ) ? 64u : 0u;
    *inb = in;
}

// Neighbour e of t when it is known to be inside the volume (every emitted
// candidate is): only the in-tile coordinate decides whether a tile is crossed.
__host__ __device__ inline uint32_t neighbor_in(const Geom& g, uint32_t t, uint32_t e)
{
    const uint32_t m = t & 4095u, hi = t - m;
    switch (e) {
    case 0: return (m & kMz) ? hi | morton_dec(m, kMz) : t - g.sTZ + kMz;
    case 1: return (m & kMy) ? hi | morton_dec(m, kMy) : t - g.sTY + kMy;
    case 2: return (m & kMx) ? hi | morton_dec(m, kMx) : t - 4096u + kMx;
    case 4: return (m & kMx) != kMx ? hi | morton_inc(m, kMx) : t + 4096u - kMx;
    case 5: return (m & kMy) != kMy ? hi | morton_inc(m, kMy) : t + g.sTY - kMy;
    case 6: return (m & kMz) != kMz ? hi | morton_inc(m, kMz) : t + g.sTZ - kMz;
    default: return t;
    }
}

__host__ __device__ inline void decode_xyz(const Geom& g, uint32_t t, uint32_t* x, uint32_t* y, uint32_t* z)
{
    const uint32_t tile = t >> 12;
    const uint32_t qy = fdiv(tile, g.dTX);
    const uint32_t tx = tile - qy * g.TX;
    const uint32_t tz = fdiv(tile, g.dTXY);
    const uint32_t ty = qy - tz * g.TY;
    const uint32_t m = t & 4095u;
    *x = tx * 16u + compact4(m);
    *y = ty * 16u + compact4(m >> 1);
    *z = tz * 16u + compact4(m >> 2);
}

__host__ __device__ inline uint64_t align256(uint64_t n) { return n + ((256u - (n % 256u)) % 256u); }

// ShouldOptimizeForFewActiveVoxels (Segmenter.cpp:1363-1369)
__host__ __device__ inline bool path_is_few(uint64_t n, uint64_t N) { return 7u * align256(n) < N; }

// ---------------------------------------------------------------------------
// device-resident control block, parameters, kernel arguments
// ---------------------------------------------------------------------------
struct DevCtl {
    uint32_t n[2];        // entries of list[0], list[1]
    uint32_t iter;        // statistics index within the current lsseg_iterate call
    uint32_t epoch;       // +1 per keep and compaction pass; stamps the look-back states
    uint32_t stats_cap;
    uint32_t error;
    uint32_t few_iters, many_iters;
    uint32_t stamp;       // the marking's byte value in C (1..255); +1 per compaction
    uint32_t pad_;
};

struct DevParams {
    float negInv;         // -1 / eps (eps truncated to an integer for int sources)
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

// Candidate emissions of the iteration being marked are counted in
// kCountSlots counters kCountStride words (128 bytes) apart, workgroup b adding
// to counter b % kCountSlots: device-scope atomics on one address serialise
// (about 11 ns each on MI355X), so a single counter would cost microseconds
// per launch.  The keep step sums and clears them.
constexpr int kCountSlots = 64;
constexpr int kCountStride = 16;
constexpr size_t kDirCounters = 2 * 7 * kCountSlots * kCountStride;  // words of KArgs::dirc

struct KArgs {
    Geom g;
    uint32_t run_log2;    // log2 of the list positions per XCD run (0: plain grid stride)
    uint32_t fused_compact_max;  // k_update_fused: at most this many workgroups (a multiple of 8) take
                                 // the compaction, the rest of its grid the update
    void* phi[2];         // float, or int8_t in the fixed-point storage mode
    uint8_t* moved;       // temporal-derivative flag per traversal position, written by the update
    const void* src;
    int8_t* E;            // export byte (tiled)
    int8_t* frozen;       // frozen export (tiled)
    uint64_t* C;          // emission record
    uint32_t* list[2];    // active lists in the reference's order
    uint32_t* S[2];       // the same voxels in a spatially coherent order (the traversal)
    uint32_t* pos;        // list position of every voxel of the current list
    uint8_t* emit;        // emission mask per traversal position of the current list
    uint8_t* keep;        // kept candidate lists of every list position
    uint8_t* memb;        // members of the next list, for a dense compaction (zero at rest)
    uint64_t* tstate;     // look-back chunk states (aggregates)
    uint64_t* gstate;     // look-back group states (totals of 64 chunks)
    uint64_t* cpart;      // candidate-emission counters (kCountSlots, zero between iterations)
    uint32_t* dirc;       // kept candidates per direction of list 0 and list 1 (7 x kCountSlots
                          // counters each, written by the keep step, cleared by the compaction)
    DevCtl* ctl;
    const DevParams* prm;
    uint32_t* n_hist;
    uint64_t* c_hist;
    uint32_t* m_hist;
    uint32_t* kcnt;       // kept voxels per chunk of a deferred keep step (LSSEG_DEFER_KEEP)
    uint32_t keep_grid;   // the keep step's grid and positions per thread (for the chunks of a
    uint32_t keep_items;  // deferred keep step); defer: the compaction writes the next count
    uint32_t defer;
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline uint64_t lanemask_lt(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

// Scans through DPP lane moves, which stay in the vector ALU (a shuffle goes
// through the LDS crossbar, ~10x the latency per step).  Every lane of the
// wave must be active.
//
// inclusive prefix within each row of 16 lanes: partial sums of 4 from the
// original value (row_shr 1, 2, 3), then of 8 and 16 (row_shr 4 on banks 1-3,
// row_shr 8 on banks 2-3)
__device__ inline uint32_t row_inclusive_scan(uint32_t v)
{
    const int x = (int)v;
    int s = x + __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
    s += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
    s += __builtin_amdgcn_update_dpp(0, x, 0x113, 0xf, 0xf, true);
    s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xf, 0xe, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xf, 0xc, false);
    return (uint32_t)s;
}

// inclusive prefix over the wave: the row scans, then row 0's total into
// rows 1 and 3 (row_bcast 15) and rows 0-1's into rows 2 and 3 (row_bcast 31)
__device__ inline uint32_t wave_inclusive_scan(uint32_t v, int /*lane*/ = 0)
{
    int s = (int)row_inclusive_scan(v);
    s += __builtin_amdgcn_update_dpp(0, s, 0x142, 0xa, 0xf, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x143, 0xc, 0xf, false);
    return (uint32_t)s;
}

// the wave's total, in every lane
__device__ inline uint32_t wave_sum(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_inclusive_scan(v), 63);
}

__device__ inline uint64_t ld_state(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void st_state(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back state: [epoch:30][flag:2][value:32]; flag 1 = a chunk's aggregate,
// 2 = a group's total.  A state written by an earlier launch carries another
// epoch and reads as "not yet published", so the states need no clearing.  The
// count travels inside the 64-bit word, so no other memory has to be ordered
// with it.
__device__ inline uint32_t state_flag(uint64_t s, uint32_t ep)
{
    return ((uint32_t)(s >> 34) == ep) ? (uint32_t)((s >> 32) & 3u) : 0u;
}

// Polls this lane's C states at words[0..C-1] (when `need`) until every lane's
// are published.  The wait is bounded: past kSpinLimit polls it sets kErrSpin,
// zeroes the values and returns.
template <int C>
__device__ __attribute__((always_inline)) inline void wait_states(const uint64_t* words, bool need, uint32_t ep,
                                                                  uint64_t* s, uint32_t& spins, uint32_t* err)
{
    // one exit (the exit test is wave-uniform): two loops with early returns
    // in one function made the compiler keep 256 VGPRs and spill
    bool waiting = true;
    while (waiting) {
        bool pending = false;
#pragma unroll
        for (int c = 0; c < C; ++c) pending |= need && state_flag(s[c], ep) == 0;
        if (__ballot(pending) == 0) {
            waiting = false;
        } else if (++spins > kSpinLimit) {
            if ((threadIdx.x & 63) == 0) atomicOr(err, kErrSpin);
#pragma unroll
            for (int c = 0; c < C; ++c) s[c] = 0;
            waiting = false;
        } else {
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (need && state_flag(s[c], ep) == 0) s[c] = ld_state(&words[c]);
        }
    }
}

// Two-level look-back over C independent counts at once, run by one wave.
// Chunk j (group g = j / 64, rank r = j % 64) publishes its aggregates tot[c]
// (agg[j * C + c], flag 1); the last chunk of a group (r = 63) also publishes
// the group's totals (grp[g * C + c], flag 2) as soon as it has read the
// group's other aggregates.  Chunk j's exclusive prefix is the sum of the
// totals of the earlier groups and of the aggregates of chunks 64g..j-1,
// written to pre[c].  With kPrevDirect the previous group is read through
// its chunks' aggregates instead of its total, which spares the chunks of
// group 1 the wait for group 0's total at the price of 64 C more loads per
// chunk (measured: without it the phantom run is 2.4% slower, the vessels
// run unchanged, for seven counts as for one).  Every chunk
// resolves after at most two rounds of publication,
// whatever its index (a chained look-back, where inclusive prefixes travel
// one 64-chunk window per round, left chunk 400 of a 428-chunk compaction
// waiting ~17 us), and waits only on chunks of lower workgroups.  Values are
// 32-bit counts.
// count c of a chunk's P published words (P = C: one count per word; P < C:
// two 16-bit counts per word)
template <int C, int P>
__device__ inline uint32_t chunk_count(const uint64_t* s, int c)
{
    if constexpr (P == C) return (uint32_t)s[c];
    else return (uint32_t)(s[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
}

// With P = (C + 1) / 2 a chunk publishes its C counts two to a word, which
// needs every count below 2^16; likewise a group's totals with G = (C + 1) / 2.
template <int C, bool kPrevDirect, int P = C, int G = C>
__device__ __attribute__((always_inline)) inline void wave_group_scan(uint64_t* agg, uint64_t* grp, uint32_t chunk, const uint32_t* tot, uint32_t epoch,
                                uint32_t* err, uint32_t* pre)
{
    static_assert(P == C || P == (C + 1) / 2, "one or two counts per word");
    static_assert(G == C || G == (C + 1) / 2, "one or two totals per word");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t ep = epoch & 0x3FFFFFFFu;
    const uint64_t stamp = (uint64_t)ep << 34;
    const uint32_t g = chunk >> 6, r = chunk & 63u;
    uint32_t word = 0;
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (P == C ? lane == (uint32_t)c : lane == (uint32_t)(c >> 1)) word |= tot[c] << (P == C ? 0 : 16 * (c & 1));
    if (lane < (uint32_t)P) st_state(&agg[(uint64_t)chunk * P + lane], stamp | (1ull << 32) | word);
    // this group's earlier chunks, the previous group's chunks and the totals
    // of the 64 groups before that, loaded together
    const bool need_own = lane < r, need_prev = kPrevDirect && g > 0;
    const uint32_t ng = kPrevDirect && g > 0 ? g - 1u : g;  // groups read through their totals
    const bool need_grp = lane < ng;
    const uint64_t* ow = agg + ((uint64_t)g * 64u + lane) * P;
    const uint64_t* pw = g ? agg + ((uint64_t)(g - 1u) * 64u + lane) * P : agg;
    const uint64_t* gw = grp + (uint64_t)lane * G;
    uint64_t so[P], sp[P], sg[G];
#pragma unroll
    for (int c = 0; c < P; ++c) {
        so[c] = need_own ? ld_state(ow + c) : 0ull;
        sp[c] = need_prev ? ld_state(pw + c) : 0ull;
    }
#pragma unroll
    for (int c = 0; c < G; ++c) sg[c] = need_grp ? ld_state(gw + c) : 0ull;
    uint32_t spins = 0;
    wait_states<P>(ow, need_own, ep, so, spins, err);
    uint32_t own[C], prior[C];
#pragma unroll
    for (int c = 0; c < C; ++c) own[c] = wave_sum(need_own ? chunk_count<C, P>(so, c) : 0u);
    if (r == 63u) {
        uint32_t gword = 0;
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (G == C ? lane == (uint32_t)c : lane == (uint32_t)(c >> 1))
                gword |= (own[c] + tot[c]) << (G == C ? 0 : 16 * (c & 1));
        if (lane < (uint32_t)G) st_state(&grp[(uint64_t)g * G + lane], stamp | (2ull << 32) | gword);
    }
    wait_states<P>(pw, need_prev, ep, sp, spins, err);
#pragma unroll
    for (int c = 0; c < C; ++c) prior[c] = wave_sum(need_prev ? chunk_count<C, P>(sp, c) : 0u);
    // the totals of groups 0..ng-1, 64 groups per round
    for (uint32_t g0 = 0; g0 < ng; g0 += 64) {
        const bool need = g0 + lane < ng;
        const uint64_t* w = grp + ((uint64_t)g0 + lane) * G;
        if (g0) {
#pragma unroll
            for (int c = 0; c < G; ++c) sg[c] = need ? ld_state(w + c) : 0ull;
        }
        wait_states<G>(w, need, ep, sg, spins, err);
#pragma unroll
        for (int c = 0; c < C; ++c) prior[c] += wave_sum(need ? chunk_count<C, G>(sg, c) : 0u);
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (lane == (uint32_t)c) pre[c] = prior[c] + own[c];
}

// the look-back of one count for a whole block: wave 0 resolves it, every
// thread gets the exclusive prefix (through *s_excl)
__device__ inline uint32_t lookback(const KArgs& a, uint32_t chunk, uint32_t total, uint32_t epoch, uint32_t* s_excl)
{
    if (threadIdx.x < 64) wave_group_scan<1, true>(a.tstate, a.gstate, chunk, &total, epoch, &a.ctl->error, s_excl);
    __syncthreads();
    return *s_excl;
}

// ---------------------------------------------------------------------------
// Look-back scans with one look-back tile per workgroup.  A scan of `items`
// sub-tiles over `participants` workgroups groups them into chunks of r
// consecutive sub-tiles, r the smallest that leaves at most `participants`
// chunks, and workgroup b takes chunk b.  A chunk waits only on chunks of lower
// workgroups, which are dispatched before it and publish their aggregates
// without waiting, so every wait ends however many workgroups are resident at
// once.  With r > 1 a chunk counts its sub-tiles, resolves its prefix, then
// places them in a second pass over their inputs.
// ---------------------------------------------------------------------------
struct ChunkPlan {
    uint32_t r;       // sub-tiles per chunk
    uint32_t chunks;  // look-back tiles
};

__device__ inline ChunkPlan chunk_plan(uint32_t items, uint32_t participants)
{
    ChunkPlan p;
    p.r = items > participants ? (items + participants - 1) / participants : 1u;
    p.chunks = (items + p.r - 1) / p.r;
    return p;
}

// ---------------------------------------------------------------------------
// the update rule (UpdateLevelSetVolume.cu:222-342), operation order kept
// ---------------------------------------------------------------------------
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

// phi storage.  float: as computed.  int8_t (the reference's LEVEL_SET_CHAR,
// Config.hpp:72-80): stored as __float2int_rd(v * 127) (UpdateLevelSetVolume.cu:
// 316-318) and read as a normalised float, s / 127 with -128 -> -1
// (cudaReadModeNormalizedFloat, Textures.cu:12).
template <typename PhiT>
__device__ inline float phi_load(const PhiT* p, uint32_t i)
{
    if constexpr (sizeof(PhiT) == 1) {
        const int s = p[i];
        return s == -128 ? -1.0f : (float)s / 127.0f;
    } else {
        return p[i];
    }
}

template <typename PhiT>
__device__ inline void phi_store(PhiT* p, uint32_t i, float v)
{
    if constexpr (sizeof(PhiT) == 1) p[i] = export_byte(v);
    else p[i] = v;
}

template <typename PhiT>
__device__ inline PhiT* phi_buf(const KArgs& a, int i)
{
    return static_cast<PhiT*>(a.phi[i]);
}

// ComputeDensityTerm (UpdateLevelSetVolume.cu:335-342): f = (-1/eps) |I - T| + 1
template <typename SrcT, bool kFloatSrc>
__device__ inline float density_factor(const KArgs& a, const DevParams& P, uint32_t t)
{
    if constexpr (kFloatSrc) {
        const float s = static_cast<const float*>(a.src)[t];
        const float d = s - P.Tf;
        return P.negInv * fabsf(d) + 1.0f;
    } else {
        const int32_t s = (int32_t)(static_cast<const SrcT*>(a.src)[t]);
        const int32_t d = s - P.Ti;
        const int32_t ad = d < 0 ? -d : d;
        return P.negInv * (float)ad + 1.0f;
    }
}

// Decisions of one voxel update, recorded by the differentiable form so that a
// replay takes the same branches: the two clamps, the connectivity gate, a
// vanishing gradient and the side of the window's absolute value.
enum : uint32_t {
    kDecLo1 = 1u,     // u4 + delta <= -1: first clamp at -1
    kDecHi1 = 2u,     // u4 + delta >= 1: first clamp at +1
    kDecLo2 = 4u,     // rescaled value <= -1: second clamp at -1
    kDecHi2 = 8u,     // rescaled value >= 1: second clamp at +1
    kDecConn = 16u,   // connectivity term applied
    kDecFlat = 32u,   // |grad phi| == 0
    kDecBelow = 64u   // float sources: I < T (the branch of |I - T|)
};

enum DecideMode : int { kDecideNone = 0, kDecideRecord = 1, kDecideReplay = 2 };

// the update of one voxel from its 7-point neighbourhood and density factor f;
// kDecideRecord stores the decisions in *dec, kDecideReplay takes them from *dec
template <int kDecide = kDecideNone>
__device__ inline float update_core(const DevParams& P, float unz, float u1, float u3, float u4, float u5, float u7,
                                    float upz, float f, uint32_t* dec = nullptr)
{
    const float dcx = (u5 - u3) / 2.0f;
    const float dcy = (u7 - u1) / 2.0f;
    const float dcz = (upz - unz) / 2.0f;
    const float fx = u5 - u4, fy = u7 - u4, fz = upz - u4;
    const float bx = u4 - u3, by = u4 - u1, bz = u4 - unz;
    const float sdx = fx - bx, sdy = fy - by, sdz = fz - bz;
    float conn = 0.0f;
    bool gate;
    if constexpr (kDecide == kDecideReplay) {
        gate = (*dec & kDecConn) != 0;
    } else {
        gate = P.alphaPositive && fabsf(upz - unz) <= kConnTol && fabsf(unz - u1) <= kConnTol &&
               fabsf(u1 - u3) <= kConnTol && fabsf(u3 - u5) <= kConnTol && fabsf(u5 - u7) <= kConnTol &&
               fabsf(u7 - upz) <= kConnTol;
    }
    if (gate) {
        const float avg = (upz + u1 + u3 + u5 + u7 + unz) / 6.0f;
        conn = avg - u4;
    }
    const float curv = sdx + sdy + sdz;
    const float dens = P.flipDensity ? f : -f;
    const float speed = (P.oneMinusAlpha * dens) + (P.alpha * curv);
    const float g = __builtin_sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
    const float delta = ((P.dtAbs * speed * g)) + conn;
    const float x1 = u4 + delta;
    float c1, nv;
    if constexpr (kDecide == kDecideReplay) {
        c1 = (*dec & kDecLo1) ? -1.0f : ((*dec & kDecHi1) ? 1.0f : x1);
        const float x2 = ((1.0f + kRescale) * c1) + kRescale;
        nv = (*dec & kDecLo2) ? -1.0f : ((*dec & kDecHi2) ? 1.0f : x2);
    } else {
        c1 = fmin_ref(1.0f, fmax_ref(-1.0f, x1));
        const float x2 = ((1.0f + kRescale) * c1) + kRescale;
        nv = fmin_ref(1.0f, fmax_ref(-1.0f, x2));
        if constexpr (kDecide == kDecideRecord) {
            uint32_t d = 0;
            if (!(x1 > -1.0f)) d |= kDecLo1;
            else if (!(x1 < 1.0f)) d |= kDecHi1;
            if (!(x2 > -1.0f)) d |= kDecLo2;
            else if (!(x2 < 1.0f)) d |= kDecHi2;
            if (gate) d |= kDecConn;
            if (g == 0.0f) d |= kDecFlat;
            *dec = d;
        }
    }
    return nv;
}

// the id of voxel (x, y, z) of the padded volume
__host__ __device__ inline uint32_t id_of(const Geom& g, uint32_t x, uint32_t y, uint32_t z)
{
    return (((z >> 4) * g.TY + (y >> 4)) * g.TX + (x >> 4)) * 4096u |
           (spread4(x & 15u) | (spread4(y & 15u) << 1) | (spread4(z & 15u) << 2));
}


// ---------------------------------------------------------------------------
// list traversal of the update and marking kernels.  Workgroups are dispatched
// to the eight XCDs round-robin (workgroup b on XCD b % 8) and each XCD has its
// own L2, so list positions are dealt to XCDs in runs of 2^run_log2 (>= the
// workgroup size): run j goes to XCD j % 8 and neighbouring list entries, which
// gather from the same cache lines, share one L2.  The walk runs over the
// `count` workgroups from `first` (both multiples of 8; by default the grid).
// ---------------------------------------------------------------------------
struct ListWalk {
    uint32_t vt, step, xcd, lr;
    __device__ ListWalk(uint32_t run_log2, uint32_t first = 0, uint32_t count = 0)
    {
        lr = run_log2;
        const uint32_t b = blockIdx.x - first;
        const uint32_t wgs = count ? count : gridDim.x;
        if (lr) {
            xcd = b & 7u;
            vt = (b >> 3) * kBlock + threadIdx.x;
            step = (wgs >> 3) * kBlock;
        } else {
            xcd = 0;
            vt = b * kBlock + threadIdx.x;
            step = wgs * kBlock;
        }
    }
    // 64-bit so that no position near 2^32 wraps back into the list
    __device__ uint64_t pos() const
    {
        if (!lr) return vt;
        const uint32_t mask = (1u << lr) - 1u;
        return ((((uint64_t)(vt >> lr) << 3) + xcd) << lr) | (vt & mask);
    }
    __device__ void next() { vt += step; }
};

// ---------------------------------------------------------------------------
// the iteration steps.  q is the parity of the current list (list[q]) and of
// the phi read buffer (phi[q]); the write buffer and the next list have parity
// q ^ 1.  Each step is a device function over one list position or one scan
// tile; the single-volume kernels and the batched kernels (which take each
// volume's arguments from a table) share them.
// ---------------------------------------------------------------------------
// Where a recording run of the differentiable form puts iteration k's active
// list V_k (in traversal order), the decisions of its updates and the values
// of the write buffer that they overwrite: from offs[k] on, with offs[0] = 0.
// Entries beyond `cap` are not written; the host then sizes the record from
// offs and records again.
struct RecordOut {
    uint64_t* offs;
    uint64_t cap;
    uint32_t* lists;
    uint8_t* dec;
    float* undo;
};

// the undo slots of a recording run's current iteration k = ctl->iter (list
// q), by traversal position, or nullptr past the capacity
__device__ inline float* record_undo(const KArgs& a, int q, const RecordOut& r)
{
    const uint32_t k = a.ctl->iter, n = a.ctl->n[q];
    const uint64_t base = r.offs[k];
    return base + n <= r.cap ? r.undo + base : nullptr;
}

// the update of active voxel u at traversal position p.  The byte moved[p]
// holds the temporal flag in bit 0 and, with kRec (a recording run of the
// differentiable form, FP32 phi), the update's decisions in bits 1-7; the
// value of W it overwrites goes to undo[p] (when undo is not null).  kClear:
// the record C[u] is cleared (the first update of a call, after which C holds
// only stamped bytes).
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kRec = false, bool kClear = true>
__device__ inline void update_at(const KArgs& a, const DevParams& P, const PhiT* __restrict__ R,
                                 PhiT* __restrict__ W, uint32_t u, uint64_t p, float* __restrict__ undo = nullptr)
{
    uint32_t nb[7], inb, d = 0;
    neighbors(a.g, u, nb, &inb);
    const float u4 = phi_load(R, u);
    float old = 0.0f;
    if constexpr (kRec)
        if (undo) old = phi_load(W, u);
    const float nv = update_core<kRec ? kDecideRecord : kDecideNone>(
        P, phi_load(R, nb[0]), phi_load(R, nb[1]), phi_load(R, nb[2]), u4, phi_load(R, nb[4]), phi_load(R, nb[5]),
        phi_load(R, nb[6]), density_factor<SrcT, kFloatSrc>(a, P, u), &d);
    phi_store(W, u, nv);
    a.E[u] = export_byte(nv);
    if constexpr (kClear) a.C[u] = 0ull;  // flags of an initialised or injected list
    // the temporal flag |phi_W - phi_R| > 0.008 of the value as computed (before
    // a fixed-point store quantises it; UpdateLevelSetVolume.cu:327-331), in
    // traversal order for the marking
    uint32_t mv = fabsf(nv - u4) > kTemporalThreshold ? 1u : 0u;
    if constexpr (kRec) {
        if constexpr (kFloatSrc)
            if (static_cast<const float*>(a.src)[u] - P.Tf < 0.0f) d |= kDecBelow;
        mv |= d << 1;
        if (undo) undo[p] = old;
    }
    a.moved[p] = (uint8_t)mv;
}

// marking of active voxel u (UpdateActiveElements.cu:75-103) whose update set
// the temporal flag `moved`; returns its emission mask
template <typename PhiT>
__device__ inline uint32_t mark_at(const KArgs& a, bool ref, bool moved, const PhiT* __restrict__ W,
                                   uint8_t* __restrict__ Cb, uint32_t u, uint32_t& pos, uint8_t stamp)
{
    uint32_t m = 0;
    // the position and every operand are loaded together (all addresses are
    // voxels of the volume); the position then decides whether u emits
    uint32_t nb[7], inb;
    neighbors(a.g, u, nb, &inb);
    pos = a.pos[u];
    const float nv = phi_load(W, u);
    float wn[7];
#pragma unroll
    for (int e = 0; e < 7; ++e) wn[e] = phi_load(W, nb[e]);
    if (!ref || (pos & 3u) != 3u) {  // list positions p % 4 == 3: UpdateActiveElementsVolume.cu:329
        if (moved) {
#pragma unroll
            for (int e = 0; e < 7; ++e) {
                if (e == 3 || !((inb >> e) & 1u)) continue;
                if (!(fabsf(nv - wn[e]) <= kTemporalThreshold)) m |= 1u << e;
            }
            if (m) m |= 8u;
        }
        if (m) {
#pragma unroll
            for (int e = 0; e < 7; ++e)
                if ((m >> e) & 1u) Cb[(uint64_t)nb[e] * 8u + (uint32_t)e] = stamp;
        }
    }
    return m;
}

// block-wide sum of c, added to *acc by one thread; every thread calls it
__device__ inline void block_add(uint64_t* acc, uint32_t c, uint32_t* s_red)
{
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += s_red[w];
        if (tot) atomicAdd((unsigned long long*)acc, (unsigned long long)tot);
    }
    __syncthreads();
}

// per-tile state of a few-path compaction (by list positions)
struct PosShared {
    uint32_t cnt[7 * 16];   // kept candidates per direction (16 each), item and wave
    uint32_t excl[7 * 16];  // their exclusive prefix within the tile, per direction
    uint32_t tot[7];                   // the tile's kept candidates per direction
    uint32_t chain[7];                 // the chunk's, summed over its tiles
    uint32_t pre[7];                   // kept candidates of direction e in the chunks before
    uint32_t run[7];                   // list offset of the next tile's direction-e candidates
    uint32_t base[8];                  // kept candidates of the directions before e; base[7] = all
};

struct CompactShared {
    uint32_t stage[kScanTile];
    uint32_t cnt[kItems * 4];
    uint32_t excl[kItems * 4];
    uint32_t prefix, total;
    PosShared pos;
};

// one compaction launch over one volume: few path over the n list positions
// (seven candidate slots each, taken in list-major order), dense path over all
// N voxels in tiled order
struct CompactPlan {
    uint32_t n_src;
    uint32_t n_dst;   // the next list's count, set by the keep step (loaded early)
    bool dense;
    uint64_t slots;   // list positions (few path) or voxels (dense path)
    uint32_t ntiles;  // sub-tiles of kBlock * items positions or kScanTile voxels
    uint32_t items;   // few path: list positions per thread of a sub-tile (1 or kPosItems)
};

template <int kMode>
__device__ inline CompactPlan compact_plan(const Geom& g, uint32_t n_src)
{
    CompactPlan c;
    c.n_src = n_src;
    c.n_dst = 0;
    c.dense = (kMode == kInit) || !path_is_few(n_src, g.N);
    c.slots = c.dense ? g.N : (uint64_t)n_src;
    c.items = (!c.dense && n_src <= g.small_n) ? 1u : (uint32_t)kPosItems;
    const uint32_t tile = c.dense ? kScanTile : kBlock * c.items;
    c.ntiles = (uint32_t)((c.slots + tile - 1) / tile);
    return c;
}

// once the last chunk is placed (one thread).  The keep step has already set
// the count of the next list (c.n_dst); the compaction must arrive at it.
template <int kMode>
__device__ inline void compact_finish(const KArgs& a, int q, const CompactPlan& c, uint32_t m, uint32_t epoch)
{
    DevCtl* ctl = a.ctl;
    if constexpr (kMode == kInit) {
        ctl->n[q ^ 1] = m;
        ctl->n[q] = 0;
    } else if (a.defer) {
        // the keep step left the count to the compaction
        ctl->n[q ^ 1] = m;
        const uint32_t k = ctl->iter - 1u;
        if (k < ctl->stats_cap) a.m_hist[k] = m;
    } else if (c.n_dst != m) {
        atomicOr(&ctl->error, kErrCount);
    }
    if constexpr (kMode == kCompact) ctl->stamp = ctl->stamp + 1u;  // no marking reads it in this launch
    ctl->epoch = epoch + 1;
}

// The compaction of list[q] into list[q ^ 1] (the reference's order): few
// path, the kept candidates of the keep bytes in list-major order (by list
// positions, compact_pos_*); dense path, the voxels marked in `memb` in tiled
// order (the marks are cleared when the voxels are placed); kInit, the voxels
// with C != 0 in tiled order.  The positions of the new list are written too.
// Every thread of the block runs these.
//
// compact_load (dense path and kInit): the flags (ballots) of sub-tile `sub`
// (voxel sub * kScanTile + j * kBlock + tid) and per-wave counts into sh.cnt;
// with kPlace also the target voxels (and the dense path's marks are cleared).
template <int kMode, bool kPlace>
__device__ __attribute__((always_inline)) inline void compact_load(const KArgs& a, int q, const CompactPlan& c,
                                                                   uint32_t sub, uint64_t* bal, uint32_t* val,
                                                                   CompactShared& sh)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)sub * kScanTile;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint64_t s = base + (uint64_t)j * kBlock + threadIdx.x;
        bool keep = false;
        uint32_t v = 0;
        if (s < c.slots) {
            v = id_of_tau((uint32_t)s);  // s runs in the reference's tiled order
            if constexpr (kMode == kInit) {
                keep = a.C[v] != 0ull;
            } else {
                keep = a.memb[v] != 0;
                if (kPlace && keep) a.memb[v] = 0;
            }
        }
        bal[j] = __ballot(keep);
        val[j] = v;
        if (lane == 0) sh.cnt[j * 4 + wave] = (uint32_t)__popcll(bal[j]);
    }
}

// block scan of a loaded sub-tile's wave counts (sh.cnt -> sh.excl); returns
// the sub-tile's total in every thread
__device__ __attribute__((always_inline)) inline uint32_t compact_counts(CompactShared& sh)
{
    const int lane = threadIdx.x & 63;
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t cc = lane < kItems * 4 ? sh.cnt[lane] : 0u;
        const uint32_t incl = wave_inclusive_scan(cc, lane);
        if (lane < kItems * 4) sh.excl[lane] = incl - cc;
        if (lane == 63) sh.total = incl;
    }
    __syncthreads();
    return sh.total;
}

// places the kept voxels of a loaded and counted sub-tile from list position
// `prefix` on, through the LDS stage so that the stores are contiguous
template <int kMode>
__device__ __attribute__((always_inline)) inline void compact_place(const KArgs& a, int q, const uint64_t* bal,
                                                                    const uint32_t* val, uint32_t prefix,
                                                                    uint32_t total, CompactShared& sh)
{
    uint32_t* __restrict__ out = a.list[q ^ 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt(lane);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if ((bal[j] >> lane) & 1ull) {
            const uint32_t pos = sh.excl[j * 4 + wave] + (uint32_t)__popcll(bal[j] & lt);
            sh.stage[pos] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < total; i += kBlock) {
        const uint32_t v = sh.stage[i];
        out[prefix + i] = v;
        a.pos[v] = prefix + i;
        if constexpr (kMode == kInit) a.S[q ^ 1][prefix + i] = v;  // tiled order is coherent
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// few-path compaction by list positions.  The next list is the concatenation
// over directions e of the kept candidates (e, p) in position order
// (Cuda.cpp:699-758 over the seven candidate lists), so candidate (e, p) goes
// to base_e + (kept e-candidates at positions before p), base_e being the kept
// candidates of the directions before e, which the keep step counts.  A tile
// of kPosTile positions reads each keep byte and list entry once and runs
// seven look-back counts, one per direction (wave_group_scan<7>).  Its inputs do
// not depend on the count, so the first workgroups load them with it.
// ---------------------------------------------------------------------------
__device__ inline uint32_t* dir_counter(const KArgs& a, int q, int e, uint32_t slot)
{
    return a.dirc + ((uint32_t)(q * 7 + e) * kCountSlots + slot) * kCountStride;
}

// this lane's direction counters of list q (one per direction; run by a wave)
__device__ inline void dir_load(const KArgs& a, int q, uint32_t* d)
{
    static_assert(kCountSlots == 64, "one counter slot per lane");
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int e = 0; e < 7; ++e) d[e] = *dir_counter(a, q, e, lane);
}

__device__ inline void dir_clear(const KArgs& a, int q)
{
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int e = 0; e < 7; ++e) *dir_counter(a, q, e, lane) = 0;
}

// keep bytes and list entries of position tile `tile` (kBlock * items
// positions, items <= kPosItems) of list[q]; positions are bounded by N (every
// one a valid address), and compact_pos_mask drops the ones at or beyond the
// count and the items past `items`
__device__ __attribute__((always_inline)) inline void compact_pos_load(const KArgs& a, int q, uint32_t tile,
                                                                       uint32_t items, uint32_t* kb, uint32_t* vx)
{
    const uint64_t base = (uint64_t)tile * kBlock * items;
#pragma unroll
    for (int j = 0; j < kPosItems; ++j)
        if (j < (int)items) {
            const uint64_t p = base + (uint64_t)j * kBlock + threadIdx.x;
            const uint64_t pc = p < a.g.N ? p : 0;
            kb[j] = a.keep[pc];
            vx[j] = a.list[q][pc];
        }
}

__device__ __attribute__((always_inline)) inline void compact_pos_mask(uint32_t n, uint32_t tile, uint32_t items,
                                                                       uint32_t* kb)
{
    const uint64_t base = (uint64_t)tile * kBlock * items;
#pragma unroll
    for (int j = 0; j < kPosItems; ++j)
        if (j >= (int)items || base + (uint64_t)j * kBlock + threadIdx.x >= n) kb[j] = 0;
}

// base_e from the direction counters d (each lane's, see dir_load), into
// ps.base (run by wave 0)
__device__ inline void compact_pos_bases(const uint32_t* d, PosShared& ps)
{
    uint32_t run = 0;
#pragma unroll
    for (int e = 0; e < 7; ++e) {
        const uint32_t t = wave_sum(d[e]);
        if (threadIdx.x == 0) ps.base[e] = run;
        run += t;
    }
    if (threadIdx.x == 0) ps.base[7] = run;
}

// per-direction counts of a loaded tile in position order (item, wave, lane):
// ps.excl, ps.tot
__device__ __attribute__((always_inline)) inline void compact_pos_counts(const uint32_t* kb, uint32_t items,
                                                                         PosShared& ps)
{
    constexpr int kSeg = 16;  // counts per direction: items (at most 4) x waves, the rest zero
    static_assert(kPosItems <= 4, "the counts of a direction fill one row of 16 lanes");
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < 7; ++e)
#pragma unroll
        for (int j = 0; j < kPosItems; ++j)
            if (j < (int)items) {
                const uint32_t c = (uint32_t)__popcll(__ballot((kb[j] >> e) & 1u));
                if (lane == 0) ps.cnt[e * 16 + j * 4 + wave] = c;
            }
    __syncthreads();
    if (threadIdx.x < 64) {
#pragma unroll
        for (int r = 0; r < (7 * kSeg + 63) / 64; ++r) {
            const int i = r * 64 + lane;
            const uint32_t cc = (i < 7 * kSeg && (uint32_t)((i % kSeg) / 4) < items) ? ps.cnt[i] : 0u;
            const uint32_t v = row_inclusive_scan(cc);
            if (i < 7 * kSeg) {
                ps.excl[i] = v - cc;
                if (i % kSeg == kSeg - 1) ps.tot[i / kSeg] = v;
            }
        }
    }
    __syncthreads();
}

// places the kept candidates of loaded and counted tile `tile`, direction e
// from list offset ps.run[e] on, and clears the tile's non-zero keep bytes
// (the keep step stores only non-zero ones)
__device__ __attribute__((always_inline)) inline void compact_pos_place(const KArgs& a, int q, uint32_t tile,
                                                                        uint32_t items, const uint32_t* kb,
                                                                        const uint32_t* vx, const PosShared& ps)
{
    const uint64_t tbase = (uint64_t)tile * kBlock * items;
#pragma unroll
    for (int j = 0; j < kPosItems; ++j)
        if (j < (int)items && kb[j]) a.keep[tbase + (uint64_t)j * kBlock + threadIdx.x] = 0;
    uint32_t* __restrict__ out = a.list[q ^ 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt(lane);
#pragma unroll
    for (int e = 0; e < 7; ++e) {
        const uint32_t r0 = ps.run[e];
#pragma unroll
        for (int j = 0; j < kPosItems; ++j) {
            if (j >= (int)items) break;
            const bool k = ((kb[j] >> e) & 1u) != 0;
            const uint64_t bal = __ballot(k);
            if (k) {
                const uint32_t at = r0 + ps.excl[e * 16 + j * 4 + wave] + (uint32_t)__popcll(bal & lt);
                const uint32_t v = neighbor_in(a.g, vx[j], (uint32_t)e);
                out[at] = v;
                a.pos[v] = at;
            }
        }
    }
}

// Look-back chunk `chunk` of a few-path compaction.  When `loaded`, kb and vx
// hold its first tile's inputs and (wave 0) d its direction counters.  Every
// chunk reads the counters before it publishes its aggregates, so the last
// chunk, whose prefixes need every other chunk's aggregates, clears them.
// Returns false (with the error flag set) if the next list would overflow.
template <int kMode>
__device__ inline bool compact_pos_chunk(const KArgs& a, int q, const CompactPlan& c, const ChunkPlan& cp,
                                         uint32_t chunk, uint32_t epoch, CompactShared& sh, uint32_t* kb,
                                         uint32_t* vx, uint32_t* d, bool loaded)
{
    PosShared& ps = sh.pos;
    const uint32_t s0 = chunk * cp.r, s1 = min(s0 + cp.r, c.ntiles);
    const bool last = chunk == cp.chunks - 1;
    if (!loaded) {
        if (threadIdx.x < 64) dir_load(a, q, d);
        if (cp.r == 1) compact_pos_load(a, q, s0, c.items, kb, vx);
    }
    if (threadIdx.x < 64) compact_pos_bases(d, ps);  // before the barriers below
    if (cp.r == 1) {
        compact_pos_mask(c.n_src, s0, c.items, kb);
        compact_pos_counts(kb, c.items, ps);
        if (threadIdx.x < 7) ps.chain[threadIdx.x] = ps.tot[threadIdx.x];
    } else {
        uint32_t acc = 0;
        for (uint32_t sub = s0; sub < s1; ++sub) {
            compact_pos_load(a, q, sub, c.items, kb, vx);
            compact_pos_mask(c.n_src, sub, c.items, kb);
            compact_pos_counts(kb, c.items, ps);
            if (threadIdx.x < 7) acc += ps.tot[threadIdx.x];
        }
        if (threadIdx.x < 7) ps.chain[threadIdx.x] = acc;
    }
    __syncthreads();
    // a chunk's count per direction is at most its positions: two counts to a
    // state word when that is below 2^16 (the same choice for every chunk)
    if (threadIdx.x < 64) {
        const uint64_t span = (uint64_t)cp.r * kBlock * c.items;  // positions per chunk
        if (a.g.pack_counts && 64u * span < 65536u)
            wave_group_scan<7, true, 4, 4>(a.tstate, a.gstate, chunk, ps.chain, epoch, &a.ctl->error, ps.pre);
        else if (a.g.pack_counts && span < 65536u)
            wave_group_scan<7, true, 4, 7>(a.tstate, a.gstate, chunk, ps.chain, epoch, &a.ctl->error, ps.pre);
        else
            wave_group_scan<7, true, 7, 7>(a.tstate, a.gstate, chunk, ps.chain, epoch, &a.ctl->error, ps.pre);
    }
    __syncthreads();
    bool over = false;
#pragma unroll
    for (int e = 0; e < 7; ++e) over |= (uint64_t)ps.base[e] + ps.pre[e] + ps.chain[e] > a.g.N;
    if (over) {
        // the last chunk overflows whenever any does (its prefixes are the
        // largest); it still advances the epoch so that no later scan reads
        // this one's states as its own
        if (threadIdx.x == 0) {
            atomicOr(&a.ctl->error, kErrOverflow);
            if (last) a.ctl->epoch = epoch + 1;
        }
        return false;
    }
    if (threadIdx.x < 7) ps.run[threadIdx.x] = ps.base[threadIdx.x] + ps.pre[threadIdx.x];
    __syncthreads();
    if (cp.r == 1) {
        compact_pos_place(a, q, s0, c.items, kb, vx, ps);
    } else {
        for (uint32_t sub = s0; sub < s1; ++sub) {
            compact_pos_load(a, q, sub, c.items, kb, vx);
            compact_pos_mask(c.n_src, sub, c.items, kb);
            compact_pos_counts(kb, c.items, ps);
            compact_pos_place(a, q, sub, c.items, kb, vx, ps);
            __syncthreads();
            if (threadIdx.x < 7) ps.run[threadIdx.x] += ps.tot[threadIdx.x];
            __syncthreads();
        }
    }
    if (last && threadIdx.x < 64) {
        // every chain ends at its direction's count from the keep step
        bool bad = false;
#pragma unroll
        for (int e = 0; e < 7; ++e) bad |= ps.pre[e] + ps.chain[e] != ps.base[e + 1] - ps.base[e];
        if (threadIdx.x == 0 && bad) atomicOr(&a.ctl->error, kErrCount);
        dir_clear(a, q);
        if (threadIdx.x == 0) compact_finish<kMode>(a, q, c, ps.base[7], epoch);
    }
    __syncthreads();
    return true;
}

// Look-back chunk `chunk` of the compaction (a few-path chunk takes the
// loaded inputs, see compact_pos_chunk).  Returns false (with the error flag
// set) if the next list would overflow, which a consistent record never does.
template <int kMode>
__device__ inline bool compact_chunk(const KArgs& a, int q, const CompactPlan& c, const ChunkPlan& cp, uint32_t chunk,
                                     uint32_t epoch, CompactShared& sh, uint32_t* kb, uint32_t* vx, uint32_t* d,
                                     bool loaded)
{
    if constexpr (kMode == kCompact)
        if (!c.dense) return compact_pos_chunk<kMode>(a, q, c, cp, chunk, epoch, sh, kb, vx, d, loaded);
    const uint32_t s0 = chunk * cp.r, s1 = min(s0 + cp.r, c.ntiles);
    uint64_t bal[kItems];
    uint32_t val[kItems];
    uint32_t total = 0;
    if (cp.r == 1) {
        compact_load<kMode, true>(a, q, c, s0, bal, val, sh);
        total = compact_counts(sh);
    } else {
        for (uint32_t sub = s0; sub < s1; ++sub) {
            compact_load<kMode, false>(a, q, c, sub, bal, val, sh);
            total += compact_counts(sh);
        }
    }
    const uint32_t prefix = lookback(a, chunk, total, epoch, &sh.prefix);
    if ((uint64_t)prefix + total > a.g.N) {
        if (threadIdx.x == 0) {
            atomicOr(&a.ctl->error, kErrOverflow);
            if (chunk == cp.chunks - 1) a.ctl->epoch = epoch + 1;  // as in compact_pos_chunk
        }
        return false;
    }
    if (cp.r == 1) {
        compact_place<kMode>(a, q, bal, val, prefix, total, sh);
    } else {
        uint32_t run = prefix;
        for (uint32_t sub = s0; sub < s1; ++sub) {
            compact_load<kMode, true>(a, q, c, sub, bal, val, sh);
            const uint32_t t = compact_counts(sh);
            compact_place<kMode>(a, q, bal, val, run, t, sh);
            run += t;
        }
    }
    if (chunk == cp.chunks - 1) {
        // the keep step's direction counters are not needed on the dense path
        if (kMode == kCompact && threadIdx.x < 64) dir_clear(a, q);
        if (threadIdx.x == 0) compact_finish<kMode>(a, q, c, prefix + total, epoch);
    }
    __syncthreads();
    return true;
}

// the whole compaction of list[q] over the first `participants` workgroups
// (every one of them calls it); `loaded` as in compact_pos_chunk, for tile
// blockIdx.x
template <int kMode>
__device__ inline void compact_all(const KArgs& a, int q, const CompactPlan& c, uint32_t epoch, CompactShared& sh,
                                   uint32_t participants, uint32_t* kb, uint32_t* vx, uint32_t* d, bool loaded)
{
    const ChunkPlan cp = chunk_plan(c.ntiles, participants);
    if (blockIdx.x < cp.chunks)
        (void)compact_chunk<kMode>(a, q, c, cp, blockIdx.x, epoch, sh, kb, vx, d, loaded && cp.r == 1);
    if (c.ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) compact_finish<kMode>(a, q, c, 0, epoch);
}

// the first workgroups' inputs of a few-path compaction of list[q], loaded
// before the counts are known (see compact_pos_chunk): the tile of kPosTile
// positions and, when small tiles are enabled, the one of kBlock positions.
// Once the plan is known compact_pick_preload selects the one it uses.
__device__ __attribute__((always_inline)) inline bool compact_preload(const KArgs& a, int q, uint32_t* kb,
                                                                      uint32_t* vx, uint32_t* d, uint32_t* small)
{
    if (blockIdx.x >= a.g.spec_tiles) return false;
    compact_pos_load(a, q, blockIdx.x, kPosItems, kb, vx);
    if (a.g.small_n) {
        const uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
        const uint64_t pc = p < a.g.N ? p : 0;
        small[0] = a.keep[pc];
        small[1] = a.list[q][pc];
    }
    if (threadIdx.x < 64) dir_load(a, q, d);
    return true;
}

__device__ __attribute__((always_inline)) inline bool compact_pick_preload(const CompactPlan& c, bool loaded,
                                                                           uint32_t* kb, uint32_t* vx,
                                                                           const uint32_t* small)
{
    if (loaded && !c.dense && c.items == 1) {
        kb[0] = small[0];
        vx[0] = small[1];
    }
    return loaded;
}

// ---------------------------------------------------------------------------
// keep step.  A sub-tile covers kKeepTile positions of S (slot j * kBlock + tid
// of the sub-tile).  Candidate (e, u -> v) is kept iff the lowest set byte of
// C[v] is e (first-wins over the list-major concatenation,
// FilterDuplicates.cu:42-66); the kept lists of u go to keep[pos[u]] for the
// compaction, and the kept voxels are appended to the next S in S order, each
// voxel's in list order: consecutive lanes write neighbouring voxels, so the
// next S is as coherent as this one.  Every voxel of the next list is kept
// exactly once; when the compaction takes the dense path it is also marked in
// `memb`.  The last chunk sets the count of the next list and the statistics.
// ---------------------------------------------------------------------------
template <int kI>
struct KeepShared {
    uint32_t cnt[kI * 4];
    uint32_t excl[kI * 4];
    uint32_t dir[2][7 * 4];  // kept candidates per direction and wave, by sub-tile parity
    uint32_t prefix, total;
};

template <int kI>
__device__ inline uint32_t keep_tiles(uint32_t n) { return (n + kBlock * kI - 1) / (kBlock * kI); }

// this lane's share of the candidate counters (run by the 64 lanes of a wave)
__device__ inline uint64_t candidate_share(const KArgs& a)
{
    uint64_t c = 0;
    for (int i = threadIdx.x & 63; i < kCountSlots; i += 64) c += a.cpart[i * kCountStride];
    return c;
}

// the control fields that keep_finish updates, read before the chunk's work
// (only keep_finish changes them during the keep step)
struct CtlSnapshot {
    uint32_t iter, stats_cap, few, many;
};

__device__ inline CtlSnapshot ctl_snapshot(const DevCtl* ctl)
{
    return {ctl->iter, ctl->stats_cap, ctl->few_iters, ctl->many_iters};
}

// counts and statistics of the iteration from list[q] (n entries) to the next
// list (m entries); c: each lane's candidate_share and st: the control fields,
// both loaded earlier; the counters are cleared.  Run by the 64 lanes of one
// wave.
__device__ inline void keep_finish(const KArgs& a, int q, uint32_t n, uint32_t m, uint32_t epoch, uint64_t c,
                                   const CtlSnapshot& st)
{
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < kCountSlots; i += 64) a.cpart[i * kCountStride] = 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += (uint64_t)__shfl_xor((unsigned long long)c, o, 64);
    if (lane != 0) return;
    DevCtl* ctl = a.ctl;
    ctl->n[q ^ 1] = m;
    const uint32_t k = st.iter;
    if (k < st.stats_cap) {
        a.n_hist[k] = n;
        a.c_hist[k] = c;
        a.m_hist[k] = m;
    }
    if (path_is_few(n, a.g.N)) ctl->few_iters = st.few + 1;
    else ctl->many_iters = st.many + 1;
    ctl->iter = k + 1;
    ctl->epoch = epoch + 1;
}

// the traversal entries of keep sub-tile `sub` of list q and their emission
// masks, with ps their list positions as a deferring marking left them in
// S[q ^ 1] (see mark_walk); positions are bounded by N (every one a valid
// address), and keep_decide discards the entries at or beyond the count
template <int kI>
__device__ __attribute__((always_inline)) inline void keep_voxels(const KArgs& a, int q, uint32_t sub, uint32_t* us,
                                                                  uint32_t* ms, uint32_t* ps = nullptr)
{
    const uint64_t base = (uint64_t)sub * (kBlock * kI);
#pragma unroll
    for (int j = 0; j < kI; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        const uint64_t ic = i < a.g.N ? i : 0;
        us[j] = a.S[q][ic];
        ms[j] = a.emit[ic];
        if (ps) ps[j] = a.S[q ^ 1][ic];
    }
}

// the kept directions kb of the voxels us (emission masks ms) of keep sub-tile
// `sub`, written to keep[pos[u]] as well (the positions psin when given, else
// gathered); entries past the count become voxel 0 with kb = 0
template <int kI>
__device__ __attribute__((always_inline)) inline void keep_decide(const KArgs& a, uint32_t n, bool dense,
                                                                  uint32_t sub, uint32_t* us, const uint32_t* ms,
                                                                  uint32_t* kb, uint32_t stamp,
                                                                  const uint32_t* psin = nullptr)
{
    const uint64_t* __restrict__ Ck = a.C;
    const uint64_t ones = 0x0101010101010101ull, st8 = ones * stamp;
    const uint64_t base = (uint64_t)sub * (kBlock * kI);
    // the positions and the records of the emitted candidates together (other
    // lanes load the voxel's own record and ignore it)
    uint32_t ps[kI];
#pragma unroll
    for (int j = 0; j < kI; ++j)
        if (base + (uint64_t)j * kBlock + threadIdx.x >= n) us[j] = 0;
#pragma unroll
    for (int j = 0; j < kI; ++j) ps[j] = psin ? psin[j] : a.pos[us[j]];
    // a -z candidate (e = 0) comes first in any record, so only e >= 1 needs one
    uint64_t rec[kI][7];
#pragma unroll
    for (int j = 0; j < kI; ++j) {
        const bool valid = base + (uint64_t)j * kBlock + threadIdx.x < n;
        const uint32_t m = valid ? ms[j] : 0u;
#pragma unroll
        for (int e = 1; e < 7; ++e) rec[j][e] = Ck[((m >> e) & 1u) ? neighbor_in(a.g, us[j], (uint32_t)e) : us[j]];
    }
#pragma unroll
    for (int j = 0; j < kI; ++j) {
        const bool valid = base + (uint64_t)j * kBlock + threadIdx.x < n;
        const uint32_t m = valid ? ms[j] : 0u;
        uint32_t k = m & 1u;
#pragma unroll
        for (int e = 1; e < 7; ++e) {
            // the bytes equal to the stamp are the zero bytes of x; the lowest
            // one is flagged exactly (no borrow reaches it from below)
            const uint64_t x = rec[j][e] ^ st8;
            const uint64_t z = (x - ones) & ~x & (ones << 7);
            if (((m >> e) & 1u) && (uint32_t)((__ffsll((unsigned long long)z) - 1) >> 3) == (uint32_t)e)
                k |= 1u << e;
        }
        kb[j] = k;
        // the keep bytes are zero between compactions (each few-path
        // compaction clears the ones it read), so only kept lists are stored,
        // and none for a dense compaction, which reads the marks instead
        if (valid && k && !dense) a.keep[ps[j]] = (uint8_t)k;
    }
}

// offsets of the kept voxels of a sub-tile: ex within the wave, sh.excl per
// wave and item; returns the sub-tile's total in every thread
template <int kI>
__device__ __attribute__((always_inline)) inline uint32_t keep_counts(const uint32_t* kb, uint32_t* ex,
                                                                      KeepShared<kI>& sh)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kI; ++j) {
        const uint32_t c = (uint32_t)__popc(kb[j]);
        const uint32_t incl = wave_inclusive_scan(c, lane);
        ex[j] = incl - c;
        if (lane == 63) sh.cnt[j * 4 + wave] = incl;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t cc = lane < kI * 4 ? sh.cnt[lane] : 0u;
        const uint32_t incl = wave_inclusive_scan(cc, lane);
        if (lane < kI * 4) sh.excl[lane] = incl - cc;
        if (lane == 63) sh.total = incl;
    }
    __syncthreads();
    return sh.total;
}

// The kept candidates of a sub-tile per direction go to this workgroup's
// direction counters of list q (the compaction's list-major bases):
// keep_dirs_count before keep_counts (whose barriers publish sh.dir),
// keep_dirs_add after it.  Buffer `buf` alternates between sub-tiles: a wave
// may count the next sub-tile while wave 0 still adds this one's.
template <int kI>
__device__ __attribute__((always_inline)) inline void keep_dirs_count(const uint32_t* kb, KeepShared<kI>& sh, int buf)
{
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < 7; ++e) {
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < kI; ++j) c += (uint32_t)__popcll(__ballot((kb[j] >> e) & 1u));
        if ((threadIdx.x & 63) == 0) sh.dir[buf][e * 4 + wave] = c;
    }
}

template <int kI>
__device__ __attribute__((always_inline)) inline void keep_dirs_add(const KArgs& a, int q, const KeepShared<kI>& sh,
                                                                    int buf)
{
    if (threadIdx.x < 7) {
        const uint32_t* w = &sh.dir[buf][threadIdx.x * 4];
        const uint32_t t = w[0] + w[1] + w[2] + w[3];
        if (t) atomicAdd(dir_counter(a, q, (int)threadIdx.x, blockIdx.x % kCountSlots), t);
    }
}

// appends the kept voxels of a counted sub-tile to the next S from `prefix` on
template <int kI>
__device__ __attribute__((always_inline)) inline void keep_append(const KArgs& a, int q, bool dense,
                                                                  const uint32_t* us, const uint32_t* kb,
                                                                  const uint32_t* ex, uint32_t prefix,
                                                                  KeepShared<kI>& sh)
{
    uint32_t* __restrict__ Sn = a.S[q ^ 1];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kI; ++j) {
        uint32_t off = prefix + sh.excl[j * 4 + wave] + ex[j];
        const uint32_t k = kb[j];
        if (k) {
#pragma unroll
            for (int e = 0; e < 7; ++e)
                if ((k >> e) & 1u) {
                    const uint32_t v = neighbor_in(a.g, us[j], (uint32_t)e);
                    Sn[off++] = v;
                    if (dense) a.memb[v] = 1;
                }
        }
    }
    __syncthreads();
}

// Look-back chunk `chunk` of the keep step; `us` and `ms` hold the entries and
// emission masks of its first sub-tile.  With several sub-tiles the first pass
// replaces each emission mask by the kept directions for the second.  Returns
// false (with the error flag set) if the next list would overflow.
template <int kI>
__device__ __attribute__((always_inline)) inline bool keep_chunk(const KArgs& a, int q, uint32_t n,
                                                                 const ChunkPlan& cp, uint32_t chunk,
                                                                 uint32_t epoch, uint32_t stamp, KeepShared<kI>& sh,
                                                                 uint32_t* us, uint32_t* ms)
{
    const bool dense = !path_is_few(n, a.g.N);
    const uint32_t s0 = chunk * cp.r, s1 = min(s0 + cp.r, keep_tiles<kI>(n));
    const bool last = chunk == cp.chunks - 1;
    // the last chunk loads the candidate counters and the control fields now
    const uint64_t cshare = (last && threadIdx.x < 64) ? candidate_share(a) : 0ull;
    const CtlSnapshot snap = last ? ctl_snapshot(a.ctl) : CtlSnapshot{0, 0, 0, 0};
    uint32_t kb[kI], ex[kI];
    uint32_t total = 0;
    if (cp.r == 1) {
        keep_decide<kI>(a, n, dense, s0, us, ms, kb, stamp);
        keep_dirs_count(kb, sh, 0);
        total = keep_counts(kb, ex, sh);
        keep_dirs_add(a, q, sh, 0);
    } else {
        for (uint32_t sub = s0; sub < s1; ++sub) {
            keep_voxels<kI>(a, q, sub, us, ms);
            keep_decide<kI>(a, n, dense, sub, us, ms, kb, stamp);
            const uint64_t base = (uint64_t)sub * (kBlock * kI);
#pragma unroll
            for (int j = 0; j < kI; ++j) {
                const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
                if (i < n) a.emit[i] = (uint8_t)kb[j];  // read back by the second pass
            }
            keep_dirs_count(kb, sh, (int)(sub & 1u));
            total += keep_counts(kb, ex, sh);
            keep_dirs_add(a, q, sh, (int)(sub & 1u));
        }
    }
    const uint32_t prefix = lookback(a, chunk, total, epoch, &sh.prefix);
    if ((uint64_t)prefix + total > a.g.N) {
        if (threadIdx.x == 0) {
            atomicOr(&a.ctl->error, kErrOverflow);
            if (last) a.ctl->epoch = epoch + 1;  // as in compact_pos_chunk
        }
        return false;
    }
    if (cp.r == 1) {
        keep_append(a, q, dense, us, kb, ex, prefix, sh);
    } else {
        uint32_t run = prefix;
        for (uint32_t sub = s0; sub < s1; ++sub) {
            keep_voxels<kI>(a, q, sub, us, ms);  // ms: the kept directions of the first pass
            const uint64_t base = (uint64_t)sub * (kBlock * kI);
#pragma unroll
            for (int j = 0; j < kI; ++j) kb[j] = base + (uint64_t)j * kBlock + threadIdx.x < n ? ms[j] : 0u;
            const uint32_t t = keep_counts(kb, ex, sh);
            keep_append(a, q, dense, us, kb, ex, run, sh);
            run += t;
        }
    }
    if (threadIdx.x < 64 && last) keep_finish(a, q, n, prefix + total, epoch, cshare, snap);
    __syncthreads();
    return true;
}

// the keep step at kI traversal positions per thread (one by default, two with
// LSSEG_KEEP_ITEMS=2), one look-back chunk per workgroup: a chunk of several
// sub-tiles when sub-tiles outnumber workgroups
template <int kI>
__global__ __launch_bounds__(kBlock) void k_keep(KArgs a, int q)
{
    __shared__ KeepShared<kI> sh;
    uint32_t us[kI], ms[kI];
    keep_voxels<kI>(a, q, blockIdx.x, us, ms);  // with the count: this workgroup's sub-tile when chunks are single
    const uint32_t epoch = a.ctl->epoch, stamp = a.ctl->stamp;
    const uint32_t n = a.ctl->n[q];
    const ChunkPlan cp = chunk_plan(keep_tiles<kI>(n), gridDim.x);
    if (blockIdx.x < cp.chunks) (void)keep_chunk<kI>(a, q, n, cp, blockIdx.x, epoch, stamp, sh, us, ms);
    if (cp.chunks == 0 && blockIdx.x == 0 && threadIdx.x < 64)
        keep_finish(a, q, 0, 0, epoch, candidate_share(a), ctl_snapshot(a.ctl));
}

// ---------------------------------------------------------------------------
// single-volume iteration kernels.  An iteration over list[q] is the update,
// the marking, the keep step and the compaction into list[q ^ 1].  The
// compaction reads only the keep bytes, the membership marks and the list, and
// writes only the next list and the positions, none of which the next
// iteration's update touches (it walks the next S, which the keep step wrote):
// k_update_fused runs the pending compaction and the next update in one launch.
// ---------------------------------------------------------------------------
// The voxels S[p] of list positions p < n, first voxel included: it is loaded
// together with the count n, not after it (S has N entries, so a position
// below N is a valid address whatever n is).
template <typename F>
__device__ inline void walk_list(const KArgs& a, int q, uint32_t first, uint32_t count, F f)
{
    const uint32_t* __restrict__ S = a.S[q];
    ListWalk w(a.run_log2, first, count);
    uint64_t p = w.pos();
    uint32_t u = S[p < a.g.N ? p : 0];
    const uint32_t n = a.ctl->n[q];
    while (p < n) {
        f(u, p);
        w.next();
        p = w.pos();
        if (p < n) u = S[p];
    }
}

template <typename SrcT, bool kFloatSrc, typename PhiT, bool kRec = false, bool kClear = true>
__device__ inline void update_walk(const KArgs& a, int q, uint32_t first = 0, uint32_t count = 0,
                                   float* undo = nullptr)
{
    const DevParams P = *a.prm;
    walk_list(a, q, first, count, [&](uint32_t u, uint64_t p) {
        update_at<SrcT, kFloatSrc, PhiT, kRec, kClear>(a, P, phi_buf<PhiT>(a, q), phi_buf<PhiT>(a, q ^ 1), u, p,
                                                       undo);
    });
}

template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update(KArgs a, int q)
{
    update_walk<SrcT, kFloatSrc, PhiT>(a, q);
}

// the update of a recording run's first iteration (FP32 phi)
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_update_record(KArgs a, int q, RecordOut r)
{
    update_walk<SrcT, kFloatSrc, float, true>(a, q, 0, 0, record_undo(a, q, r));
}

// the compaction of list[q ^ 1] into list[q] and the update over S[q].  The
// first workgroups (a multiple of 8, at most three quarters of the grid and
// no more than the compaction has sub-tiles) take the compaction, the others
// the update, so neither waits behind the other.  The grid is a multiple of 8
// and at least 16.
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kRec>
__device__ inline void update_fused_body(const KArgs& a, int q, const RecordOut& r)
{
    __shared__ CompactShared sh;
    uint32_t kb[kPosItems], vx[kPosItems], d[7], small[2] = {0u, 0u};
    bool loaded = compact_preload(a, q ^ 1, kb, vx, d, small);  // with the counts
    const uint32_t epoch = a.ctl->epoch;
    CompactPlan c = compact_plan<kCompact>(a.g, a.ctl->n[q ^ 1]);
    c.n_dst = a.ctl->n[q];
    loaded = compact_pick_preload(c, loaded, kb, vx, small);
    const uint32_t gc = min((c.ntiles + 7u) & ~7u, a.fused_compact_max);
    if (blockIdx.x < gc) {
        compact_all<kCompact>(a, q ^ 1, c, epoch, sh, gc, kb, vx, d, loaded);
    } else {
        if (c.ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) compact_finish<kCompact>(a, q ^ 1, c, 0, epoch);
        float* undo = nullptr;
        if constexpr (kRec) undo = record_undo(a, q, r);
        update_walk<SrcT, kFloatSrc, PhiT, kRec, false>(a, q, gc, gridDim.x - gc, undo);
    }
}

template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update_fused(KArgs a, int q)
{
    update_fused_body<SrcT, kFloatSrc, PhiT, false>(a, q, RecordOut{});
}

// k_update_fused of a recording run (FP32 phi)
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_update_fused_record(KArgs a, int q, RecordOut r)
{
    update_fused_body<SrcT, kFloatSrc, float, true>(a, q, r);
}

// The counts and statistics of `count` iterations over an empty list, which
// change nothing else (one workgroup): both lists stay empty, the iterations
// take the few path.
__global__ __launch_bounds__(kBlock) void k_empty_iterations(KArgs a, uint32_t count)
{
    DevCtl* ctl = a.ctl;
    const uint32_t k0 = ctl->iter, cap = ctl->stats_cap;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < count; i += kBlock)
        if (k0 + i < cap) {
            a.n_hist[k0 + i] = 0;
            a.c_hist[k0 + i] = 0;
            a.m_hist[k0 + i] = 0;
        }
    if (threadIdx.x == 0) {
        ctl->n[0] = 0;
        ctl->n[1] = 0;
        ctl->few_iters += count;
        ctl->iter = k0 + count;
    }
}

// The marking over S[q]; with kRec also the recording of the iteration, k =
// ctl->iter: V_k in traversal order from offs[k] on, with the decisions that
// the update left in bits 1-7 of moved[p], and offs[k + 1].  The replay and the
// adjoint take V_k in any order; the traversal order keeps their gathers and
// atomics as coherent as the forward's, and the recording's stores coalesced.
template <typename PhiT, bool kRec>
__device__ inline void mark_walk(const KArgs& a, int q, const RecordOut& r)
{
    __shared__ uint32_t s_red[kBlock / 64];
    const bool ref = a.prm->refMode != 0;
    uint64_t base = 0;
    bool rec = false;
    if constexpr (kRec) {
        const uint32_t k = a.ctl->iter, n = a.ctl->n[q];
        base = r.offs[k];
        if (blockIdx.x == 0 && threadIdx.x == 0) r.offs[k + 1] = base + n;
        rec = base + n <= r.cap;
    }
    // a deferred keep step reads the list positions in traversal order from
    // S[q ^ 1], which is free until the next launch rebuilds the traversal
    uint32_t* __restrict__ sp = a.defer ? a.S[q ^ 1] : nullptr;
    const uint32_t stamp32 = a.ctl->stamp;
    const uint8_t stamp = (uint8_t)stamp32;
    if (stamp32 - 1u > 254u && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.ctl->error, kErrStamp);
    uint32_t emitted = 0;
    walk_list(a, q, 0, 0, [&](uint32_t u, uint64_t p) {
        const uint32_t mv = a.moved[p];
        uint32_t pos;
        const uint32_t m =
            mark_at<PhiT>(a, ref, (mv & 1u) != 0, phi_buf<PhiT>(a, q ^ 1), (uint8_t*)a.C, u, pos, stamp);
        a.emit[p] = (uint8_t)m;
        if (sp) sp[p] = pos;
        emitted += (uint32_t)__popc(m);
        if constexpr (kRec)
            if (rec) {
                r.lists[base + p] = u;
                r.dec[base + p] = (uint8_t)(mv >> 1);
            }
    });
    block_add(&a.cpart[(blockIdx.x % kCountSlots) * kCountStride], emitted, s_red);
}

template <typename PhiT>
__global__ __launch_bounds__(kBlock) void k_mark(KArgs a, int q)
{
    mark_walk<PhiT, false>(a, q, RecordOut{});
}

__global__ __launch_bounds__(kBlock) void k_mark_record(KArgs a, int q, RecordOut r)
{
    mark_walk<float, true>(a, q, r);
}

template <int kMode>
__global__ __launch_bounds__(kBlock) void k_compact(KArgs a, int q)
{
    __shared__ CompactShared sh;
    uint32_t kb[kPosItems], vx[kPosItems], d[7], small[2] = {0u, 0u};
    bool loaded = kMode == kCompact && compact_preload(a, q, kb, vx, d, small);
    const uint32_t epoch = a.ctl->epoch;
    CompactPlan c = compact_plan<kMode>(a.g, a.ctl->n[q]);
    c.n_dst = a.ctl->n[q ^ 1];
    loaded = compact_pick_preload(c, loaded, kb, vx, small);
    compact_all<kMode>(a, q, c, epoch, sh, gridDim.x, kb, vx, d, loaded);
}

// ---------------------------------------------------------------------------
// The keep step without a look-back (LSSEG_DEFER_KEEP, single volumes).  The
// next traversal's order is free: the update, the marking and the keep step
// give the same results whatever order they walk the list in, and only the
// compaction follows the reference's order.  So the keep step does not place
// the kept voxels: it leaves each traversal position's kept directions in
// emit[p] and its chunk's count in kcnt, and stops.  The update launch that
// follows rebuilds the next traversal from them (the kept targets of
// positions in order, so exactly the S the keep step would have appended) as
// it updates: a chunk's start is the sum of the counts before it, summed
// after the launch boundary with plain loads instead of a cross-XCD
// publication and poll.  The compaction in the same launch writes the next
// list's count, which it knows from the direction counters.
// ---------------------------------------------------------------------------

// counts and statistics of a deferred keep step of list q (n entries): as
// keep_finish, but the next count is written by the compaction and the
// look-back epoch is untouched
__device__ inline void keep_finish_defer(const KArgs& a, uint32_t n, uint64_t c, const CtlSnapshot& st)
{
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < kCountSlots; i += 64) a.cpart[i * kCountStride] = 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += (uint64_t)__shfl_xor((unsigned long long)c, o, 64);
    if (lane != 0) return;
    DevCtl* ctl = a.ctl;
    const uint32_t k = st.iter;
    if (k < st.stats_cap) {
        a.n_hist[k] = n;
        a.c_hist[k] = c;
    }
    if (path_is_few(n, a.g.N)) ctl->few_iters = st.few + 1;
    else ctl->many_iters = st.many + 1;
    ctl->iter = k + 1;
}

// chunk `chunk` of a deferred keep step: decisions (keep bytes, kept
// directions into emit, dense marks), direction counters and the chunk's count
template <int kI>
__device__ __attribute__((always_inline)) inline void keep_defer_chunk(const KArgs& a, int q, uint32_t n,
                                                                      const ChunkPlan& cp, uint32_t chunk,
                                                                      uint32_t stamp, KeepShared<kI>& sh,
                                                                      uint32_t* us, uint32_t* ms, uint32_t* ps)
{
    const bool dense = !path_is_few(n, a.g.N);
    const uint32_t s0 = chunk * cp.r, s1 = min(s0 + cp.r, keep_tiles<kI>(n));
    const bool last = chunk == cp.chunks - 1;
    const uint64_t cshare = (last && threadIdx.x < 64) ? candidate_share(a) : 0ull;
    const CtlSnapshot snap = last ? ctl_snapshot(a.ctl) : CtlSnapshot{0, 0, 0, 0};
    uint32_t kb[kI], ex[kI];
    uint32_t total = 0;
    for (uint32_t sub = s0; sub < s1; ++sub) {
        if (cp.r > 1) keep_voxels<kI>(a, q, sub, us, ms, ps);
        keep_decide<kI>(a, n, dense, sub, us, ms, kb, stamp, ps);
        const uint64_t base = (uint64_t)sub * (kBlock * kI);
#pragma unroll
        for (int j = 0; j < kI; ++j) {
            const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
            if (i < n) a.emit[i] = (uint8_t)kb[j];
            if (dense && kb[j]) {
#pragma unroll
                for (int e = 0; e < 7; ++e)
                    if ((kb[j] >> e) & 1u) a.memb[neighbor_in(a.g, us[j], (uint32_t)e)] = 1;
            }
        }
        keep_dirs_count(kb, sh, (int)(sub & 1u));
        total += keep_counts(kb, ex, sh);
        keep_dirs_add(a, q, sh, (int)(sub & 1u));
    }
    if (threadIdx.x == 0) a.kcnt[chunk] = total;
    if (threadIdx.x < 64 && last) keep_finish_defer(a, n, cshare, snap);
}

// workgroups of a deferred keep step that load their sub-tile speculatively
// (the resident slots of the whole chip at full occupancy)
constexpr uint32_t kKeepSpec = 2048;

template <int kI>
__global__ __launch_bounds__(kBlock) void k_keep_defer(KArgs a, int q)
{
    __shared__ KeepShared<kI> sh;
    uint32_t us[kI], ms[kI], ps[kI];
    // the first kKeepSpec workgroups load their sub-tile with the count (its
    // chunk when chunks are single); a wider grid's others load it once they
    // know they have one, so idle workgroups on short lists load nothing
    const bool spec = blockIdx.x < kKeepSpec;
    if (spec) keep_voxels<kI>(a, q, blockIdx.x, us, ms, ps);
    const uint32_t n = a.ctl->n[q], stamp = a.ctl->stamp;
    const ChunkPlan cp = chunk_plan(keep_tiles<kI>(n), gridDim.x);
    if (blockIdx.x < cp.chunks) {
        if (!spec && cp.r == 1) keep_voxels<kI>(a, q, blockIdx.x, us, ms, ps);
        keep_defer_chunk<kI>(a, q, n, cp, blockIdx.x, stamp, sh, us, ms, ps);
    }
    if (cp.chunks == 0 && blockIdx.x == 0 && threadIdx.x < 64) keep_finish_defer(a, 0, candidate_share(a), ctl_snapshot(a.ctl));
}

struct AppendShared {
    uint32_t tgt[7 * kBlock];  // kept targets of one block of positions, in order
    uint32_t cnt[4];           // per-wave counts
    uint32_t excl[4];
    uint32_t red[4];
    uint32_t total;
};

union FusedDeferShared {
    CompactShared c;
    AppendShared d;
};

// sum over the block of v (every thread gets it; two barriers)
__device__ __attribute__((always_inline)) inline uint32_t block_sum(uint32_t v, uint32_t* red)
{
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Positions [p0, p1) of the traversal S[q ^ 1] expanded into their kept
// targets (kept directions in emit), stored into S[q] from position `run` on
// in order, each voxel updated as it is placed when kUpd.  Every thread of the
// block calls it.
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kUpd, bool kRec = false>
__device__ __attribute__((always_inline)) inline void append_chunk(const KArgs& a, int q, const DevParams& P,
                                                                  uint64_t p0, uint64_t p1, uint32_t run,
                                                                  AppendShared& sh, float* undo = nullptr,
                                                                  uint64_t undo_room = 0)
{
    const uint32_t* __restrict__ Ss = a.S[q ^ 1];
    uint32_t* __restrict__ Sn = a.S[q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t blk = p0; blk < p1; blk += kBlock) {
        const uint64_t p = blk + threadIdx.x;
        const bool valid = p < p1;
        const uint32_t u = valid ? Ss[p] : 0u;
        const uint32_t kb = valid ? a.emit[p] : 0u;
        const uint32_t c = (uint32_t)__popc(kb);
        const uint32_t incl = wave_inclusive_scan(c, lane);
        if (lane == 63) sh.cnt[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t c0 = sh.cnt[0], c1 = sh.cnt[1], c2 = sh.cnt[2], c3 = sh.cnt[3];
            sh.excl[0] = 0;
            sh.excl[1] = c0;
            sh.excl[2] = c0 + c1;
            sh.excl[3] = c0 + c1 + c2;
            sh.total = c0 + c1 + c2 + c3;
        }
        __syncthreads();
        uint32_t off = sh.excl[wave] + incl - c;
#pragma unroll
        for (int e = 0; e < 7; ++e)
            if ((kb >> e) & 1u) sh.tgt[off++] = neighbor_in(a.g, u, (uint32_t)e);
        const uint32_t t = sh.total;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < t; i += kBlock) {
            const uint32_t v = sh.tgt[i];
            Sn[run + i] = v;
            if constexpr (kRec)
                update_at<SrcT, kFloatSrc, PhiT, true, false>(a, P, phi_buf<PhiT>(a, q), phi_buf<PhiT>(a, q ^ 1), v,
                                                              run + i, run + i < undo_room ? undo : nullptr);
            else if constexpr (kUpd)
                update_at<SrcT, kFloatSrc, PhiT, false, false>(a, P, phi_buf<PhiT>(a, q), phi_buf<PhiT>(a, q ^ 1), v,
                                                               run + i);
        }
        run += t;
        __syncthreads();
    }
}

// The traversal of list q rebuilt from a deferred keep step over list q ^ 1
// (its traversal S[q ^ 1], kept directions in emit, chunk counts in kcnt),
// over `count` workgroups from `first` (multiples of 8), each voxel updated as
// it is placed when kUpd.  Chunks go to the XCDs in runs of 16 (4096 positions
// of single-tile chunks), as ListWalk deals list positions.
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kUpd, bool kRec = false>
__device__ inline void append_walk(const KArgs& a, int q, uint32_t first, uint32_t count, AppendShared& sh,
                                   const RecordOut& r = RecordOut{})
{
    const uint32_t nsrc = a.ctl->n[q ^ 1];
    const uint32_t items = (uint32_t)(kBlock * a.keep_items);
    const ChunkPlan cp = chunk_plan((nsrc + items - 1) / items, a.keep_grid);
    const uint64_t span = (uint64_t)cp.r * items;  // positions per chunk
    DevParams P{};
    if constexpr (kUpd) P = *a.prm;
    // a recording run's undo slots of this iteration (k = ctl->iter) from
    // offs[k] on, as many as fit below the capacity: the list's count is
    // placed in this launch, so each entry checks its own slot
    float* undo = nullptr;
    uint64_t undo_room = 0;
    if constexpr (kRec) {
        const uint64_t base = r.offs[a.ctl->iter];
        undo = r.undo + base;
        undo_room = base < r.cap ? r.cap - base : 0;
    }
    const uint32_t b = blockIdx.x - first, g8 = count >> 3, x = b & 7u;
    for (uint32_t s = b >> 3;; s += g8) {
        // the s-th chunk of XCD x's runs; j grows with s
        const uint32_t j = (((s >> 4) << 3) + x) * 16u + (s & 15u);
        if (j >= cp.chunks) break;
        // the chunk's start: the counts of the chunks before it
        uint32_t pre = 0;
        for (uint32_t i = threadIdx.x; i < j; i += kBlock) pre += a.kcnt[i];
        const uint32_t run = block_sum(pre, sh.red);
        const uint64_t p0 = (uint64_t)j * span;
        append_chunk<SrcT, kFloatSrc, PhiT, kUpd, kRec>(a, q, P, p0, min(p0 + span, (uint64_t)nsrc), run, sh, undo,
                                                        undo_room);
    }
}

// the compaction of list[q ^ 1] into list[q] and the update over list q,
// whose traversal it rebuilds from the deferred keep step (see append_walk)
template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update_fused_defer(KArgs a, int q)
{
    __shared__ FusedDeferShared sh;
    uint32_t kb[kPosItems], vx[kPosItems], d[7], small[2] = {0u, 0u};
    bool loaded = compact_preload(a, q ^ 1, kb, vx, d, small);  // with the counts
    const uint32_t epoch = a.ctl->epoch;
    CompactPlan c = compact_plan<kCompact>(a.g, a.ctl->n[q ^ 1]);
    loaded = compact_pick_preload(c, loaded, kb, vx, small);
    const uint32_t gc = min((c.ntiles + 7u) & ~7u, a.fused_compact_max);
    if (blockIdx.x < gc) {
        compact_all<kCompact>(a, q ^ 1, c, epoch, sh.c, gc, kb, vx, d, loaded);
    } else {
        if (c.ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) compact_finish<kCompact>(a, q ^ 1, c, 0, epoch);
        append_walk<SrcT, kFloatSrc, PhiT, true>(a, q, gc, gridDim.x - gc, sh.d);
    }
}

// k_update_fused_defer of a recording run (FP32 phi): the update records
// each voxel's decisions and the value it overwrites by its rebuilt position
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_update_fused_defer_record(KArgs a, int q, RecordOut r)
{
    __shared__ FusedDeferShared sh;
    uint32_t kb[kPosItems], vx[kPosItems], d[7], small[2] = {0u, 0u};
    bool loaded = compact_preload(a, q ^ 1, kb, vx, d, small);  // with the counts
    const uint32_t epoch = a.ctl->epoch;
    CompactPlan c = compact_plan<kCompact>(a.g, a.ctl->n[q ^ 1]);
    loaded = compact_pick_preload(c, loaded, kb, vx, small);
    const uint32_t gc = min((c.ntiles + 7u) & ~7u, a.fused_compact_max);
    if (blockIdx.x < gc) {
        compact_all<kCompact>(a, q ^ 1, c, epoch, sh.c, gc, kb, vx, d, loaded);
    } else {
        if (c.ntiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) compact_finish<kCompact>(a, q ^ 1, c, 0, epoch);
        append_walk<SrcT, kFloatSrc, float, true, true>(a, q, gc, gridDim.x - gc, sh.d, r);
    }
}

// the traversal of list q rebuilt after a call's last deferred keep step
__global__ __launch_bounds__(kBlock) void k_settle_traversal(KArgs a, int q)
{
    __shared__ AppendShared sh;
    append_walk<uint8_t, false, float, false>(a, q, 0, gridDim.x, sh);
}

// ---------------------------------------------------------------------------
// batched iteration kernels: many volumes of uneven size, each with its own
// parameters, advanced by one launch per step.  Work items (blocks of kBlock
// list positions, or look-back tiles) are numbered volume after volume; every
// workgroup rebuilds the per-volume offsets from the device counts, so the
// launches need no host input and replay unchanged from a captured graph.
// Volume v's control block is ctls[v] (also vols[v].ctl).
// ---------------------------------------------------------------------------
constexpr uint32_t kMaxBatch = 1024;

struct BatchArgs {
    const KArgs* vols;
    DevCtl* ctls;
    uint32_t count;
    uint32_t keep_items;  // traversal positions per thread of k_keep_batch: 1 or 2 (autotuned)
    uint32_t keep_grid;   // the keep step's grid (its chunk plan, for a deferred keep step)
    uint32_t defer;       // deferred keep step: the update launch rebuilds each volume's traversal
};

// exclusive prefix s_pre[0..count] of f(v) over the volumes (all threads)
template <typename F>
__device__ inline void batch_prefix(uint32_t count, F f, uint32_t* s_pre, uint32_t* s_tmp)
{
    const uint32_t per = (count + kBlock - 1) / kBlock;  // <= kMaxBatch / kBlock
    const uint32_t v0 = threadIdx.x * per;
    uint32_t vals[kMaxBatch / kBlock];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kMaxBatch / kBlock; ++i) {
        const uint32_t v = v0 + i;
        vals[i] = (i < per && v < count) ? f(v) : 0u;
        sum += vals[i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_scan(sum, lane);
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0;
    for (int w = 0; w < wave; ++w) wbase += s_tmp[w];
    uint32_t run = wbase + incl - sum;
#pragma unroll
    for (uint32_t i = 0; i < kMaxBatch / kBlock; ++i) {
        const uint32_t v = v0 + i;
        if (i < per && v < count) s_pre[v] = run;
        run += vals[i];
    }
    if (threadIdx.x == kBlock - 1) s_pre[count] = wbase + incl;
    __syncthreads();
}

// the volume owning work item j: s_pre[v] <= j < s_pre[v + 1]
__device__ inline uint32_t batch_find(const uint32_t* s_pre, uint32_t count, uint32_t j)
{
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[mid] <= j) lo = mid;
        else hi = mid;
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
}

// the update of every volume over its S[q]: work items of kBlock positions,
// volume after volume, item j taken by slot j mod nslots (kClear: see update_at)
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kClear>
__device__ inline void update_batch_items(const BatchArgs& b, int q, uint32_t slot, uint32_t nslots, uint32_t* s_pre,
                                          uint32_t* s_tmp)
{
    batch_prefix(b.count, [&](uint32_t v) { return (b.ctls[v].n[q] + kBlock - 1) / kBlock; }, s_pre, s_tmp);
    const uint32_t items = s_pre[b.count];
    for (uint32_t j = slot; j < items; j += nslots) {
        const uint32_t v = batch_find(s_pre, b.count, j);
        const KArgs& a = b.vols[v];
        const uint64_t p = (uint64_t)(j - s_pre[v]) * kBlock + threadIdx.x;
        if (p < b.ctls[v].n[q]) {
            const DevParams P = *a.prm;
            update_at<SrcT, kFloatSrc, PhiT, false, kClear>(a, P, phi_buf<PhiT>(a, q), phi_buf<PhiT>(a, q ^ 1),
                                                            a.S[q][p], p);
        }
    }
}

template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update_batch(BatchArgs b, int q)
{
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    update_batch_items<SrcT, kFloatSrc, PhiT, true>(b, q, blockIdx.x, gridDim.x, s_pre, s_tmp);
}

template <typename PhiT>
__global__ __launch_bounds__(kBlock) void k_mark_batch(BatchArgs b, int q)
{
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    __shared__ uint32_t s_red[kBlock / 64];
    batch_prefix(b.count, [&](uint32_t v) { return (b.ctls[v].n[q] + kBlock - 1) / kBlock; }, s_pre, s_tmp);
    const uint32_t items = s_pre[b.count];
    for (uint32_t j = blockIdx.x; j < items; j += gridDim.x) {
        const uint32_t v = batch_find(s_pre, b.count, j);
        const KArgs& a = b.vols[v];
        const uint64_t p = (uint64_t)(j - s_pre[v]) * kBlock + threadIdx.x;
        uint32_t emitted = 0;
        if (threadIdx.x == 0 && b.ctls[v].stamp - 1u > 254u) atomicOr(&b.ctls[v].error, kErrStamp);
        if (p < b.ctls[v].n[q]) {
            uint32_t pos;
            const uint32_t m = mark_at<PhiT>(a, a.prm->refMode != 0, (a.moved[p] & 1u) != 0, phi_buf<PhiT>(a, q ^ 1),
                                             (uint8_t*)a.C, a.S[q][p], pos, (uint8_t)b.ctls[v].stamp);
            a.emit[p] = (uint8_t)m;
            if (b.defer) a.S[q ^ 1][p] = pos;  // for the deferred keep step (see mark_walk)
            emitted = (uint32_t)__popc(m);
        }
        block_add(&a.cpart[(blockIdx.x % kCountSlots) * kCountStride], emitted, s_red);
    }
}

// Sub-tiles per look-back chunk of a batched scan: with t_v sub-tiles in
// volume v and G workgroups, r = 1 when the sum T of the t_v is at most G, else
// ceil(T / (G - count)), which leaves at most T / r + count <= G chunks (the
// host keeps G above the volume count).  Workgroup j takes chunk j of the
// volume-major numbering.
__device__ inline uint32_t batch_chunk_r(uint32_t total, uint32_t count, uint32_t G)
{
    if (total <= G) return 1u;
    const uint32_t room = G > count ? G - count : 1u;
    return (total + room - 1) / room;
}

template <int kI>
__device__ __attribute__((always_inline)) inline void keep_batch_body(const BatchArgs& b, int q, KeepShared<kI>& sh,
                                                                      uint32_t* s_pre, uint32_t* s_tmp)
{
    batch_prefix(b.count, [&](uint32_t v) { return keep_tiles<kI>(b.ctls[v].n[q]); }, s_pre, s_tmp);
    const uint32_t r = batch_chunk_r(s_pre[b.count], b.count, gridDim.x);
    __syncthreads();
    batch_prefix(b.count, [&](uint32_t v) { return (keep_tiles<kI>(b.ctls[v].n[q]) + r - 1) / r; }, s_pre, s_tmp);
    const uint32_t j = blockIdx.x;
    if (j < s_pre[b.count]) {
        const uint32_t v = batch_find(s_pre, b.count, j);
        const KArgs& a = b.vols[v];
        const uint32_t n = b.ctls[v].n[q];
        const ChunkPlan cp{r, (keep_tiles<kI>(n) + r - 1) / r};
        const uint32_t chunk = j - s_pre[v];
        uint32_t us[kI], ms[kI], ps[kI];
        const uint32_t stamp = b.ctls[v].stamp;
        if (b.defer) {
            keep_voxels<kI>(a, q, chunk * r, us, ms, ps);
            keep_defer_chunk<kI>(a, q, n, cp, chunk, stamp, sh, us, ms, ps);
        } else {
            keep_voxels<kI>(a, q, chunk * r, us, ms);
            (void)keep_chunk<kI>(a, q, n, cp, chunk, b.ctls[v].epoch, stamp, sh, us, ms);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_keep_batch(BatchArgs b, int q)
{
    __shared__ KeepShared<1> sh1;
    __shared__ KeepShared<2> sh2;
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    if (b.keep_items == 1) keep_batch_body<1>(b, q, sh1, s_pre, s_tmp);
    else keep_batch_body<2>(b, q, sh2, s_pre, s_tmp);
    // volumes with an empty list still advance their counts
    if (threadIdx.x < 64)
        for (uint32_t v = blockIdx.x; v < b.count; v += gridDim.x)
            if (s_pre[v + 1] == s_pre[v]) {
                if (b.defer)
                    keep_finish_defer(b.vols[v], 0, candidate_share(b.vols[v]), ctl_snapshot(&b.ctls[v]));
                else
                    keep_finish(b.vols[v], q, 0, 0, b.ctls[v].epoch, candidate_share(b.vols[v]),
                                ctl_snapshot(&b.ctls[v]));
            }
}

// The traversals of list q of every volume rebuilt from a deferred batched
// keep step over list q ^ 1 (see append_walk), each voxel updated as it is
// placed when kUpd: work items are the keep step's chunks, volume after
// volume, item j taken by slot j mod nslots.
template <typename SrcT, bool kFloatSrc, typename PhiT, bool kUpd>
__device__ inline void append_batch_items(const BatchArgs& b, int q, uint32_t slot, uint32_t nslots,
                                          uint32_t* s_pre, uint32_t* s_tmp, AppendShared& sh)
{
    const uint32_t kI = b.keep_items, items = kBlock * kI;
    // the keep step's plan: r sub-tiles per chunk over all volumes' sub-tiles
    batch_prefix(b.count, [&](uint32_t v) { return (b.ctls[v].n[q ^ 1] + items - 1) / items; }, s_pre, s_tmp);
    const uint32_t r = batch_chunk_r(s_pre[b.count], b.count, b.keep_grid);
    __syncthreads();
    batch_prefix(b.count, [&](uint32_t v) { return ((b.ctls[v].n[q ^ 1] + items - 1) / items + r - 1) / r; }, s_pre,
                 s_tmp);
    const uint32_t total = s_pre[b.count];
    const uint64_t span = (uint64_t)r * items;
    for (uint32_t j = slot; j < total; j += nslots) {
        const uint32_t v = batch_find(s_pre, b.count, j);
        const KArgs& a = b.vols[v];
        const uint32_t chunk = j - s_pre[v];
        DevParams P{};
        if constexpr (kUpd) P = *a.prm;
        uint32_t pre = 0;
        for (uint32_t i = threadIdx.x; i < chunk; i += kBlock) pre += a.kcnt[i];
        const uint32_t run = block_sum(pre, sh.red);
        const uint64_t p0 = (uint64_t)chunk * span;
        append_chunk<SrcT, kFloatSrc, PhiT, kUpd>(a, q, P, p0, min(p0 + span, (uint64_t)b.ctls[v].n[q ^ 1]), run,
                                                  sh);
    }
}

// the traversals of list q of every volume rebuilt after a batched call's last
// deferred keep step
__global__ __launch_bounds__(kBlock) void k_settle_batch(BatchArgs b, int q)
{
    __shared__ AppendShared sh;
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    append_batch_items<uint8_t, false, float, false>(b, q, blockIdx.x, gridDim.x, s_pre, s_tmp, sh);
}

// The compaction of list[q] of every volume over the workgroups of slots
// [0, participants) (each calls it with its slot, blockIdx.x).  A volume's
// epoch is read at each of its chunks: its last chunk, which advances the
// epoch, completes only after every other chunk of the volume has published.
__device__ inline void compact_batch_chunks(const BatchArgs& b, int q, uint32_t participants, CompactShared& sh,
                                            uint32_t* s_pre, uint32_t* s_tmp)
{
    batch_prefix(b.count, [&](uint32_t v) { return compact_plan<kCompact>(b.vols[v].g, b.ctls[v].n[q]).ntiles; },
                 s_pre, s_tmp);
    const uint32_t r = batch_chunk_r(s_pre[b.count], b.count, participants);
    __syncthreads();
    batch_prefix(
        b.count,
        [&](uint32_t v) { return (compact_plan<kCompact>(b.vols[v].g, b.ctls[v].n[q]).ntiles + r - 1) / r; },
        s_pre, s_tmp);
    const uint32_t j = blockIdx.x;
    if (j < s_pre[b.count]) {
        const uint32_t v = batch_find(s_pre, b.count, j);
        const KArgs& a = b.vols[v];
        CompactPlan c = compact_plan<kCompact>(a.g, b.ctls[v].n[q]);
        c.n_dst = b.ctls[v].n[q ^ 1];
        const ChunkPlan cp{r, (c.ntiles + r - 1) / r};
        uint32_t kb[kPosItems], vx[kPosItems], d[7];
        (void)compact_chunk<kCompact>(a, q, c, cp, j - s_pre[v], b.ctls[v].epoch, sh, kb, vx, d, false);
    }
    // volumes with nothing to compact still advance their epochs
    if (threadIdx.x == 0)
        for (uint32_t v = blockIdx.x; v < b.count; v += participants)
            if (s_pre[v + 1] == s_pre[v]) {
                CompactPlan c = compact_plan<kCompact>(b.vols[v].g, 0);
                c.n_dst = b.ctls[v].n[q ^ 1];
                compact_finish<kCompact>(b.vols[v], q, c, 0, b.ctls[v].epoch);
            }
}

__global__ __launch_bounds__(kBlock) void k_compact_batch(BatchArgs b, int q)
{
    __shared__ CompactShared sh;
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    compact_batch_chunks(b, q, gridDim.x, sh, s_pre, s_tmp);
}

// the batched counterpart of k_update_fused: workgroups [0, gc) compact every
// volume's list[q ^ 1] into list[q], the others run the update over S[q]
template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update_fused_batch(BatchArgs b, int q, uint32_t gc)
{
    __shared__ FusedDeferShared sh;
    __shared__ uint32_t s_pre[kMaxBatch + 1];
    __shared__ uint32_t s_tmp[kBlock / 64];
    if (blockIdx.x < gc) compact_batch_chunks(b, q ^ 1, gc, sh.c, s_pre, s_tmp);
    else if (b.defer)
        append_batch_items<SrcT, kFloatSrc, PhiT, true>(b, q, blockIdx.x - gc, gridDim.x - gc, s_pre, s_tmp, sh.d);
    else
        update_batch_items<SrcT, kFloatSrc, PhiT, false>(b, q, blockIdx.x - gc, gridDim.x - gc, s_pre, s_tmp);
}

#define LSSEG_FOR_EACH(p, count)                                                           \
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < (count); p += gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------
// differentiable form.  A recording run stores, per iteration k, the active
// list V_k and the decisions of every update; a replay runs the update over the
// recorded lists with the recorded decisions (the same values as the recorded
// run for the same inputs, and a smooth function of phi0, T, eps and alpha for
// perturbed ones); the adjoint walks the iterations backwards.
//
// One iteration maps the buffer pair (A, B) (read, write) to
//   A'(v) = f_v(A) for v in V_k,  A'(v) = B(v) otherwise,  B' = A,
// so with adjoints (Abar', Bbar') of the outputs
//   Abar = Bbar' + sum_{v in V_k} J_v^T Abar'(v),  Bbar = Abar' outside V_k (0 on V_k).
// k_adjoint reads and clears X = Abar' on V_k (leaving Bbar) and scatters the
// J_v^T terms into Y = Bbar' (leaving Abar).
// ---------------------------------------------------------------------------
// One recorded update from R into W.
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_replay(KArgs a, const uint32_t* __restrict__ V,
                                                   const uint8_t* __restrict__ dec, uint32_t n,
                                                   const float* __restrict__ R, float* __restrict__ W)
{
    const DevParams P = *a.prm;
    LSSEG_FOR_EACH(p, n)
    {
        const uint32_t u = V[p];
        uint32_t nb[7], inb, d = dec[p];
        neighbors(a.g, u, nb, &inb);
        W[u] = update_core<kDecideReplay>(P, R[nb[0]], R[nb[1]], R[nb[2]], R[u], R[nb[4]], R[nb[5]], R[nb[6]],
                                          density_replay<SrcT, kFloatSrc>(a, P, u, d), &d);
    }
}

// density factor with the recorded side of |I - T| (float sources); the same
// value as density_factor at the recorded inputs, since T - I == -(I - T)
template <typename SrcT, bool kFloatSrc>
__device__ inline float density_replay(const KArgs& a, const DevParams& P, uint32_t u, uint32_t d)
{
    if constexpr (kFloatSrc) {
        const float di = static_cast<const float*>(a.src)[u] - P.Tf;
        return P.negInv * ((d & kDecBelow) ? -di : di) + 1.0f;
    } else {
        return density_factor<SrcT, kFloatSrc>(a, P, u);
    }
}

// block-wide sum of three values into theta[0..2] (FP32 atomics); every
// thread calls it.  theta is the workgroup's slot of kCountSlots slots
// kThetaStride floats (128 bytes) apart, which the host adds up.
constexpr int kThetaStride = 32;
__device__ inline void block_add3(float* theta, float v0, float v1, float v2, float* s_red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v0 += __shfl_xor(v0, o, 64);
        v1 += __shfl_xor(v1, o, 64);
        v2 += __shfl_xor(v2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_red[(threadIdx.x >> 6) * 3 + 0] = v0;
        s_red[(threadIdx.x >> 6) * 3 + 1] = v1;
        s_red[(threadIdx.x >> 6) * 3 + 2] = v2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const float t = (s_red[threadIdx.x] + s_red[3 + threadIdx.x]) + (s_red[6 + threadIdx.x] + s_red[9 + threadIdx.x]);
        if (t != 0.0f) atomicAdd(&theta[threadIdx.x], t);
    }
}

// Adjoint of iteration k: A holds the read buffer A_k; X = Abar_{k+1} (cleared
// on V_k), Y = Bbar_{k+1} (receives the J^T terms); theta[0..2] accumulates
// dL/dT, dL/deps, dL/dalpha (T and eps only for float sources: integer sources
// truncate them, Segmenter.cpp:817-818, and have zero derivative).
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_adjoint(KArgs a, const uint32_t* __restrict__ V,
                                                    const uint8_t* __restrict__ dec, uint32_t n,
                                                    const float* __restrict__ A, float* __restrict__ X,
                                                    float* __restrict__ Y, float* theta,
                                                    float* __restrict__ Bres, const float* __restrict__ undo)
{
    __shared__ float s_red[3 * kBlock / 64];
    static_assert(kBlock == 256, "block_add3 sums four waves");
    const DevParams P = *a.prm;
    const float dd = P.flipDensity ? 1.0f : -1.0f;
    float tT = 0.0f, tE = 0.0f, tA = 0.0f;
    LSSEG_FOR_EACH(p, n)
    {
        const uint32_t u = V[p];
        const uint32_t d = dec[p];
        // the write buffer of iteration k back to its values before the update
        // (kept by the recording run); this kernel reads the other buffer, A
        if (undo) Bres[u] = undo[p];
        const float gbar = X[u];
        X[u] = 0.0f;
        if ((d & (kDecLo1 | kDecHi1 | kDecLo2 | kDecHi2)) || gbar == 0.0f) continue;
        const float s = gbar * (1.0f + kRescale);  // through both clamps, inside
        uint32_t nb[7], inb;
        neighbors(a.g, u, nb, &inb);
        const float unz = A[nb[0]], u1 = A[nb[1]], u3 = A[nb[2]], u4 = A[u], u5 = A[nb[4]], u7 = A[nb[5]],
                    upz = A[nb[6]];
        const float dcx = (u5 - u3) / 2.0f, dcy = (u7 - u1) / 2.0f, dcz = (upz - unz) / 2.0f;
        const float curv = ((u5 - u4) - (u4 - u3)) + ((u7 - u4) - (u4 - u1)) + ((upz - u4) - (u4 - unz));
        const float g = __builtin_sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
        const float f = density_replay<SrcT, kFloatSrc>(a, P, u, d);
        const float dens = dd * f;
        const float speed = P.oneMinusAlpha * dens + P.alpha * curv;
        const float sp = s * P.dtAbs * g;      // dL/dspeed
        const float gs = s * P.dtAbs * speed;  // dL/d|grad phi|
        const float gc = sp * P.alpha;         // dL/dcurvature
        const float cn = (d & kDecConn) ? s / 6.0f : 0.0f;
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        if (!(d & kDecFlat) && g > 0.0f) {
            const float k = 0.5f * gs / g;
            gx = k * dcx;
            gy = k * dcy;
            gz = k * dcz;
        }
        atomicAdd(&Y[nb[0]], gc + cn - gz);
        atomicAdd(&Y[nb[1]], gc + cn - gy);
        atomicAdd(&Y[nb[2]], gc + cn - gx);
        atomicAdd(&Y[u], s - 6.0f * gc - ((d & kDecConn) ? s : 0.0f));
        atomicAdd(&Y[nb[4]], gc + cn + gx);
        atomicAdd(&Y[nb[5]], gc + cn + gy);
        atomicAdd(&Y[nb[6]], gc + cn + gz);
        tA += sp * (curv - dens);
        if constexpr (kFloatSrc) {
            // |I - T| on its recorded side: sgn (I - T)
            const float di = static_cast<const float*>(a.src)[u] - P.Tf;
            const float sgn = (d & kDecBelow) ? -1.0f : 1.0f;
            const float dspeed = sp * P.oneMinusAlpha * dd;
            tT += dspeed * (-P.negInv * sgn);
            tE += dspeed * (P.negInv * P.negInv * (sgn * di));
        }
    }
    block_add3(theta + (blockIdx.x % kCountSlots) * kThetaStride, tT, tE, tA, s_red);
}

// linear unpadded <-> tiled padded
__global__ void k_lin_to_tiled(Geom g, const float* __restrict__ lin, float* __restrict__ tiled, uint32_t nx,
                               uint32_t ny, uint32_t nz, float pad)
{
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x, y, z;
        decode_xyz(g, (uint32_t)t, &x, &y, &z);
        tiled[t] = (x < nx && y < ny && z < nz) ? lin[((uint64_t)z * ny + y) * nx + x] : pad;
    }
}

__global__ void k_tiled_sum_to_lin(Geom g, const float* __restrict__ a, const float* __restrict__ b,
                                   float* __restrict__ lin, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint64_t cnt = (uint64_t)nx * ny * nz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cnt;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % nx);
        const uint32_t y = (uint32_t)((i / nx) % ny);
        const uint32_t z = (uint32_t)(i / ((uint64_t)nx * ny));
        const uint32_t t = id_of(g, x, y, z);
        lin[i] = b ? a[t] + b[t] : a[t];
    }
}

// ---------------------------------------------------------------------------
// dense helpers (initialisation, readback)
// ---------------------------------------------------------------------------
#define LSSEG_GRID_STRIDE(i, count)                                                        \
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (count);       \
         i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_fill_f32(float* p, uint64_t count, float v)
{
    LSSEG_GRID_STRIDE(i, count) p[i] = v;
}

// random-access floor (measurement aid, not on the path): thread t < count
// makes `loads` loads (or one store) of 4 bytes, each in a pseudo-random
// 32-byte sector of the `words` 4-byte words at a; every index is reduced
// modulo the sector count, so all accesses stay in the buffer
__device__ inline uint32_t probe_mix(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

template <int kLoads, bool kStore>
__global__ __launch_bounds__(kBlock) void k_probe_access(float* a, uint64_t words, uint32_t count, uint32_t seed,
                                                         float* sink)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const uint64_t sectors = words / 8u;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kLoads; ++k) {
        const uint32_t i = t * (uint32_t)kLoads + (uint32_t)k;
        const uint64_t h = ((uint64_t)probe_mix(i ^ seed) << 32) | probe_mix(i * 2654435761u + seed);
        const uint64_t w = (h % sectors) * 8u + (i & 7u);
        if constexpr (kStore) a[w] = 0.f;
        else s += a[w];
    }
    if (!kStore && s == 12345.678f && t < words) sink[t] = s;  // keeps the loads; a zeroed buffer never takes it
}

__global__ void k_fill_u8(uint8_t* p, uint64_t count, uint8_t v)
{
    LSSEG_GRID_STRIDE(i, count) p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t count, uint64_t v)
{
    LSSEG_GRID_STRIDE(i, count) p[i] = v;
}

// every emission record cleared and the stamp restarted (before it would pass
// 255; see the emission record C): at `start`, 1, or 0 when the compaction of
// the last marking is still to run and advances it to 1
__global__ void k_records_reset(uint64_t* C, uint64_t count, DevCtl* ctl, uint32_t start)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->stamp = start;
    LSSEG_GRID_STRIDE(i, count) C[i] = 0ull;
}

// pad + tile the linear source (LoadVolumeHostPadded, Segmenter.cpp:1131-1204)
__global__ void k_pad_tile_source(Geom g, const uint8_t* lin, uint8_t* dst, uint32_t nx, uint32_t ny,
                                  uint32_t nz, uint32_t bytes, uint8_t pad)
{
    LSSEG_GRID_STRIDE(t, g.N)
    {
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

// phi0 injection: both buffers, export byte, dense band flags in C (the
// band test uses phi0 itself; fixed-point storage saturates beyond [-1, 1))
template <typename PhiT>
__global__ void k_set_phi(KArgs a, const float* lin, uint32_t nx, uint32_t ny, uint32_t nz, float band)
{
    LSSEG_GRID_STRIDE(t, a.g.N)
    {
        uint32_t x, y, z;
        decode_xyz(a.g, (uint32_t)t, &x, &y, &z);
        float v = 1.0f;
        if (x < nx && y < ny && z < nz) v = lin[((uint64_t)z * ny + y) * nx + x];
        phi_store(phi_buf<PhiT>(a, 0), (uint32_t)t, v);
        phi_store(phi_buf<PhiT>(a, 1), (uint32_t)t, v);
        a.E[t] = export_byte(v);
        a.C[t] = fabsf(v) < band ? 1ull : 0ull;
    }
}

// InitializeLevelSetKernel (InitializeLevelSetVolume.cu:55-98)
template <typename PhiT>
__global__ void k_seed_pass(KArgs a, int rd, uint32_t sx, uint32_t sy, uint32_t sz, int32_t radius, float aniso)
{
    const PhiT* R = phi_buf<PhiT>(a, rd);
    PhiT* W = phi_buf<PhiT>(a, rd ^ 1);
    LSSEG_GRID_STRIDE(t, a.g.N)
    {
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
        float nv = phi_load(R, (uint32_t)t) + c - 1.0f;
        nv = nv < -1.0f ? -1.0f : (nv > 1.0f ? 1.0f : nv);
        phi_store(W, (uint32_t)t, nv);
        a.E[t] = export_byte(nv);
    }
}

// InitializeActiveElementsVolumeKernel (InitializeActiveElementsVolume.cu:37-74):
// dense keep flags into C
template <typename PhiT>
__global__ void k_keep_flags(KArgs a, int rd)
{
    const PhiT* R = phi_buf<PhiT>(a, rd);
    LSSEG_GRID_STRIDE(t, a.g.N)
    {
        uint32_t nb[7], inb;
        neighbors(a.g, (uint32_t)t, nb, &inb);
        const float u4 = phi_load(R, (uint32_t)t);
        bool flat = true;
#pragma unroll
        for (int e = 0; e < 7; ++e) {
            if (e == 3) continue;
            if (!(fabsf(u4 - phi_load(R, nb[e])) <= kSpatialThreshold)) flat = false;
        }
        a.C[t] = flat ? 0ull : 1ull;
    }
}

// update only (the parameter-change sub-step, Segmenter.cpp:811-876)
template <typename SrcT, bool kFloatSrc, typename PhiT>
__global__ __launch_bounds__(kBlock) void k_update_only(KArgs a, int op)
{
    const DevParams P = *a.prm;
    const uint32_t n = a.ctl->n[op];
    const PhiT* R = phi_buf<PhiT>(a, op);
    PhiT* W = phi_buf<PhiT>(a, op ^ 1);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t u = a.list[op][p];
        uint32_t nb[7], inb;
        neighbors(a.g, u, nb, &inb);
        const float nv = update_core(P, phi_load(R, nb[0]), phi_load(R, nb[1]), phi_load(R, nb[2]), phi_load(R, u),
                                     phi_load(R, nb[4]), phi_load(R, nb[5]), phi_load(R, nb[6]),
                                     density_factor<SrcT, kFloatSrc>(a, P, u));
        phi_store(W, u, nv);
        a.E[u] = export_byte(nv);
    }
}

// list positions of an injected list
__global__ void k_scatter_positions(const uint32_t* list, uint32_t n, uint32_t* pos)
{
    LSSEG_GRID_STRIDE(p, n) pos[list[p]] = (uint32_t)p;
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

// CollectCurrentLevelSetVoxels / CalculateSegmentationParameters
// (Segmenter.cpp:357-478): source values of the voxels whose export byte is
// <= 0.  mode 0 sums the values and counts them, mode 1 sums the squared
// deviations from `mean`, in FP32 as the reference does: each thread with a
// compensated (Kahan) sum, then a tree over the workgroup into partial[block];
// the host adds the partials pairwise.  So only the summation order differs
// from the reference's sequential float sums, and the error stays near one
// rounding instead of growing with the count.
template <typename SrcT, bool kFloatSrc>
__global__ __launch_bounds__(kBlock) void k_inside_moments(KArgs a, int mode, float mean, float* partial,
                                                           unsigned long long* count)
{
    __shared__ float s_red[kBlock / 64];
    __shared__ unsigned long long s_cnt[kBlock / 64];
    float v = 0.0f, comp = 0.0f;
    unsigned long long c = 0;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < a.g.N;
         t += (uint64_t)gridDim.x * blockDim.x) {
        if (a.E[t] <= 0) {
            const float x = kFloatSrc ? static_cast<const float*>(a.src)[t] : (float)static_cast<const SrcT*>(a.src)[t];
            const float d = x - mean;
            const float y = (mode == 0 ? x : d * d) - comp;
            const float s = v + y;
            comp = (s - v) - y;
            v = s;
            ++c;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_red[threadIdx.x >> 6] = v;
        s_cnt[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
        const unsigned long long tc = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        if (mode == 0 && tc) atomicAdd(count, tc);
    }
}

__global__ void k_freeze(KArgs a)
{
    LSSEG_GRID_STRIDE(t, a.g.N) a.frozen[t] = add_export(a.frozen[t], a.E[t]);
}

template <typename PhiT>
__global__ void k_get_phi(KArgs a, int which, float* out, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint64_t cnt = (uint64_t)nx * ny * nz;
    LSSEG_GRID_STRIDE(i, cnt)
    {
        const uint32_t x = (uint32_t)(i % nx);
        const uint32_t y = (uint32_t)((i / nx) % ny);
        const uint32_t z = (uint32_t)(i / ((uint64_t)nx * ny));
        out[i] = phi_load(phi_buf<PhiT>(a, which), id_of(a.g, x, y, z));
    }
}

// the raw state of a fixed-point phi buffer as floats (k / 127) and back; the
// way back rounds, so a value k / 127 returns to k
__global__ void k_phi8_to_float(const int8_t* in, float* out, uint64_t count)
{
    LSSEG_GRID_STRIDE(i, count) out[i] = phi_load(in, (uint32_t)i);
}

__global__ void k_float_to_phi8(const float* in, int8_t* out, uint64_t count)
{
    LSSEG_GRID_STRIDE(i, count)
    {
        float q = rintf(in[i] * 127.0f);
        q = q > 127.0f ? 127.0f : (q < -128.0f ? -128.0f : q);
        out[i] = (int8_t)(int32_t)q;
    }
}

// Check mode (LSSEG_CHECKS): the invariants of list[q] after its compaction.
// Every entry lies inside the volume and sits at the position that pos records
// for it, so no voxel appears twice (the reference's duplicate removal relies
// on duplicate-free lists, FilterDuplicates.cu:127-193, main.tex:399); every
// traversal entry S[q][p] is a member of the list.  A violation sets kErrCheck.
__global__ __launch_bounds__(kBlock) void k_check_list(KArgs a, int q)
{
    const uint32_t n = a.ctl->n[q];
    const uint32_t* __restrict__ L = a.list[q];
    const uint32_t* __restrict__ S = a.S[q];
    bool bad = false;
    LSSEG_GRID_STRIDE(p, n)
    {
        const uint32_t v = L[p];
        bad |= v >= a.g.N || a.pos[v] != (uint32_t)p;
        const uint32_t u = S[p];
        const uint32_t pu = u < a.g.N ? a.pos[u] : 0xFFFFFFFFu;
        bad |= pu >= n || L[pu < n ? pu : 0] != u;
    }
    if (bad) atomicOr(&a.ctl->error, kErrCheck);
}

// the parameters and a fresh control block, written by the device (callers
// that stage nothing in host memory)
__global__ void k_put_params_ctl(DevParams* prm, DevParams P, DevCtl* ctl, DevCtl c0)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *prm = P;
        *ctl = c0;
    }
}

// sums of the per-iteration counters n_k, c_k, m_k of `count` iterations (one
// workgroup of kBlock threads)
struct IterSums {
    unsigned long long n, c, m;
};

__global__ __launch_bounds__(kBlock) void k_sum_history(const uint32_t* n, const uint64_t* c, const uint32_t* m,
                                                        uint32_t count, IterSums* out)
{
    __shared__ unsigned long long part[3][kBlock / 64];
    unsigned long long sn = 0, sc = 0, sm = 0;
    for (uint32_t i = threadIdx.x; i < count; i += kBlock) {
        sn += n[i];
        sc += c[i];
        sm += m[i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sn += __shfl_xor(sn, o, 64);
        sc += __shfl_xor(sc, o, 64);
        sm += __shfl_xor(sm, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = sn;
        part[1][threadIdx.x >> 6] = sc;
        part[2][threadIdx.x >> 6] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        IterSums r{0, 0, 0};
        for (int w = 0; w < kBlock / 64; ++w) {
            r.n += part[0][w];
            r.c += part[1][w];
            r.m += part[2][w];
        }
        *out = r;
    }
}

// GetSaveableSegmentation (Segmenter.cpp:1216-1286): export unioned with the
// frozen export, predicate <= 0, unpadded (UnpadVolume, :1288-1337)
__global__ void k_get_mask(KArgs a, uint8_t* out, uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint64_t cnt = (uint64_t)nx * ny * nz;
    LSSEG_GRID_STRIDE(i, cnt)
    {
        const uint32_t x = (uint32_t)(i % nx);
        const uint32_t y = (uint32_t)((i / nx) % ny);
        const uint32_t z = (uint32_t)(i / ((uint64_t)nx * ny));
        const uint32_t t = id_of(a.g, x, y, z);
        out[i] = add_export(a.E[t], a.frozen[t]) <= 0 ? 1 : 0;
    }
}

}  // namespace lsseg
