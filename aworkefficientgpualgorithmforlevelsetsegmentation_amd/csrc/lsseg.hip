// lsseg.hip -- C ABI of the sparse level-set engine for MI355X (gfx950).
//
// Implements include/lsseg/lsseg.h on top of the device code in
// lsseg_device.hpp.  Reference paths in comments are relative to
// code/gpuseg/code/ of the reference repository.
//
// An lsseg_iterate call of `iters` iterations launches, per iteration, the
// update, the marking and the compaction of the current list.  All counts stay
// in device memory; the host never reads n between iterations.  Runs of
// iterations are captured in hipGraphs.
//
// State at rest (between calls): list[q] holds the active list (count n[q]),
// phi[q] is the read buffer and the emission record C is non-zero exactly at the
// list's voxels.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lsseg/lsseg.h"
#include "lsseg_device.hpp"

using namespace lsseg;

namespace {
thread_local std::string g_create_error;
constexpr int kGraphIters = 32;  // iterations per captured graph (even: the parity returns)
constexpr int kProfileWatch = 64; // a profiled run checks for an empty list every this many iterations
}  // namespace

struct lsseg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;          // false for the volumes of a batch (shared stream)
    lsseg_batch* batch = nullptr;     // the owning batch, if any
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t nx = 0, ny = 0, nz = 0;
    int32_t dtype = 0;
    uint32_t S = 1;
    bool phi8 = false;                // fixed-point (int8) phi storage
    size_t P = 4;                     // bytes per phi value
    float z_aniso = 1.0f;
    int q = 0;                        // parity of the current list and phi read buffer
    uint32_t grid = 1024;             // co-resident workgroups of the compaction kernel
    uint32_t fgrid = 1024;            // grid of k_update_fused (compaction + update)
    uint32_t sgrid = 2048;            // workgroups of the update and marking kernels
    const void* keep_fn = nullptr;    // k_keep<1> or k_keep<2> (LSSEG_KEEP_ITEMS)
    uint32_t grid0 = 1024;            // the keep/compaction grid chosen at creation (occupancy x CUs)
    uint32_t small_n_max = 0;         // the small-tile bound at creation (the look-back states are sized for it)
    uint32_t cus = 256;
    KArgs a{};
    DevParams* d_prm = nullptr;
    DevParams h_prm{};
    DevCtl* h_ctl = nullptr;          // pinned mirror
    uint32_t* h_watch = nullptr;      // pinned: list counts after the last two graph chunks
    hipEvent_t watch_ev[2] = {nullptr, nullptr};
    uint32_t stats_cap = 0;
    uint64_t last_iters = 0;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    struct Checkpoint {
        void* phi[2] = {nullptr, nullptr};
        int8_t* E = nullptr;
        uint32_t* list = nullptr;
        uint32_t* S = nullptr;        // the traversal order of the list
        uint32_t n = 0;
        bool valid = false;
    } ckpt[LSSEG_MAX_CHECKPOINTS];
    bool profile = false;             // time every iteration launch with its own start/stop events
    bool checks = false;              // LSSEG_CHECKS: verify the list after every compaction
    bool defer_keep = false;          // LSSEG_DEFER_KEEP: the keep step without a look-back, the next
                                      // update rebuilding the traversal (single volumes)
    uint32_t keep_grid_min = 4096;    // LSSEG_KEEP_GRID: least grid of a deferred keep step
    uint32_t stamp = 0;               // the device's DevCtl::stamp at the next marking (see
                                      // stamp_reserve); 0: the records are cleared first
    std::vector<hipEvent_t> prof_events;
    size_t prof_next = 0;
    std::vector<void*> allocs;
    std::vector<lsseg_record*> records;  // live records of the differentiable form
    uint64_t record_total = 0;        // sum_k n_k of the last record, the next record's first guess
    int32_t record_iters = 0;
    std::string err;
};

namespace {

void detach_records(lsseg_ctx* c);
int env_int(const char* name, int dflt);

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

#define HIP_TRY(ctx, expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), LSSEG_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

// LSSEG_FAIL_ALLOC=k (tests of the out-of-memory paths): the k-th device
// allocation of each context creation (lsseg_create, and each volume of
// lsseg_batch_create) fails as if the device were out of memory.
thread_local int t_create_allocs = 0;  // allocations of the creation in progress (0: none in progress)

hipError_t device_malloc(void** p, size_t bytes)
{
    if (t_create_allocs > 0) {
        const int k = env_int("LSSEG_FAIL_ALLOC", 0);
        if (k > 0 && t_create_allocs++ == k) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
    }
    return hipMalloc(p, bytes);
}

struct CreateAllocScope {
    CreateAllocScope() { t_create_allocs = 1; }
    ~CreateAllocScope() { t_create_allocs = 0; }
};

lsseg_status dalloc(lsseg_ctx* c, void** p, size_t bytes)
{
    hipError_t e = device_malloc(p, bytes ? bytes : 1);
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
    for (auto& g : c->graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// LDS bytes per CU that the grid sizing counts on (LSSEG_LDS_BUDGET, default
// 64 KiB, the most one workgroup may use)
int lds_budget()
{
    const int v = env_int("LSSEG_LDS_BUDGET", 65536);
    return v >= 4096 ? v : 65536;
}

// Resident workgroups per CU that every listed kernel reaches.  The look-back
// scans hand their tiles out in order, so they finish with any grid; a grid of
// this size keeps every workgroup running from the start.  The runtime's
// figure is capped by one computed from the kernel's own resources with a
// 64 KiB LDS budget per CU and the 512-entry register file of a SIMD (one wave
// of each kBlock workgroup per SIMD, at most 8 waves per SIMD).
int min_occupancy(std::initializer_list<const void*> kernels)
{
    int best = 1 << 30;
    for (const void* k : kernels) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, kBlock, 0) != hipSuccess || occ < 1) {
            (void)hipGetLastError();
            occ = 1;
        }
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, k) == hipSuccess) {
            const int regs = std::max(8, (fa.numRegs + 7) / 8 * 8);
            const int by_regs = 512 / regs;
            const int by_lds = fa.sharedSizeBytes ? (int)(lds_budget() / fa.sharedSizeBytes) : 8;
            occ = std::min(occ, std::max(1, std::min(8, std::min(by_regs, by_lds))));
        } else {
            (void)hipGetLastError();
        }
        best = std::min(best, occ);
    }
    return best;
}

// x[0] + ... + x[n - 1] by pairwise summation in FP32, whose rounding error
// grows with log n instead of n
float pairwise_sum(const float* x, size_t n)
{
    if (n <= 8) {
        float t = 0.0f;
        for (size_t i = 0; i < n; ++i) t += x[i];
        return t;
    }
    const size_t h = n / 2;
    return pairwise_sum(x, h) + pairwise_sum(x + h, n - h);
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

// kernel instantiations for a source type and a phi storage
#define LSSEG_SRC_PHI_KERNEL(fn, name)                                                                 \
    template <typename PhiT>                                                                          \
    const void* fn##_t(int32_t dtype)                                                                 \
    {                                                                                                 \
        switch (dtype) {                                                                              \
        case LSSEG_U8: return (const void*)&name<uint8_t, false, PhiT>;                               \
        case LSSEG_I8: return (const void*)&name<int8_t, false, PhiT>;                                \
        case LSSEG_U16: return (const void*)&name<uint16_t, false, PhiT>;                             \
        case LSSEG_I16: return (const void*)&name<int16_t, false, PhiT>;                              \
        case LSSEG_U32: return (const void*)&name<uint32_t, false, PhiT>;                             \
        case LSSEG_I32: return (const void*)&name<int32_t, false, PhiT>;                              \
        default: return (const void*)&name<float, true, PhiT>;                                        \
        }                                                                                             \
    }                                                                                                 \
    const void* fn(int32_t dtype, bool phi8) { return phi8 ? fn##_t<int8_t>(dtype) : fn##_t<float>(dtype); }
LSSEG_SRC_PHI_KERNEL(update_kernel, k_update)
LSSEG_SRC_PHI_KERNEL(update_fused_kernel, k_update_fused)
LSSEG_SRC_PHI_KERNEL(update_only_kernel, k_update_only)
LSSEG_SRC_PHI_KERNEL(update_batch_kernel, k_update_batch)
LSSEG_SRC_PHI_KERNEL(update_fused_batch_kernel, k_update_fused_batch)
LSSEG_SRC_PHI_KERNEL(update_fused_defer_kernel, k_update_fused_defer)
#undef LSSEG_SRC_PHI_KERNEL

const void* mark_kernel(bool phi8) { return phi8 ? (const void*)&k_mark<int8_t> : (const void*)&k_mark<float>; }
const void* mark_batch_kernel(bool phi8)
{
    return phi8 ? (const void*)&k_mark_batch<int8_t> : (const void*)&k_mark_batch<float>;
}

// a single-volume kernel of the phi storage, launched on the context's stream
template <typename... Args>
hipError_t launch_phi(lsseg_ctx* c, void (*f32)(Args...), void (*i8)(Args...), uint32_t grid, Args... args)
{
    void* packed[] = {(void*)&args...};
    return hipLaunchKernel(c->phi8 ? (const void*)i8 : (const void*)f32, dim3(grid), dim3(kBlock), packed, 0,
                           c->stream);
}

void fill_phi_one(lsseg_ctx* c, void* buf)
{
    const uint64_t N = c->a.g.N;
    if (c->phi8)
        hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, (uint8_t*)buf, N,
                           (uint8_t)127);
    else
        hipLaunchKernelGGL(k_fill_f32, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, (float*)buf, N, 1.0f);
}

void launch(lsseg_ctx* c, const void* fn, uint32_t grid, int q)
{
    void* args[] = {(void*)&c->a, (void*)&q};
    if (c->profile && c->prof_next + 2 <= c->prof_events.size()) {
        // the events take the kernel's own start and end timestamps
        hipEvent_t e0 = c->prof_events[c->prof_next], e1 = c->prof_events[c->prof_next + 1];
        c->prof_next += 2;
        (void)hipExtLaunchKernel(fn, dim3(grid), dim3(kBlock), args, 0, c->stream, e0, e1, 0);
        return;
    }
    (void)hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, 0, c->stream);
}

// LSSEG_CHECKS: the invariants of list[q] once its compaction is done
void launch_check(lsseg_ctx* c, int q)
{
    if (c->checks) hipLaunchKernelGGL(k_check_list, dim3(c->sgrid), dim3(kBlock), 0, c->stream, c->a, q);
}

// The update over S[q]; `fused`: in the same launch as the compaction of
// list[q ^ 1] into list[q] that the previous iteration left pending (then
// checked in check mode).
void launch_head(lsseg_ctx* c, int q, bool fused)
{
    if (fused) {
        launch(c, c->a.defer ? update_fused_defer_kernel(c->dtype, c->phi8) : update_fused_kernel(c->dtype, c->phi8),
               c->fgrid, q);
        launch_check(c, q);
    } else {
        launch(c, update_kernel(c->dtype, c->phi8), c->sgrid, q);
    }
}

// marking and keep step of the iteration over list[q]; afterwards the count
// and the traversal of the next list are set, its compaction is pending
void launch_tail(lsseg_ctx* c, int q)
{
    launch(c, mark_kernel(c->phi8), c->sgrid, q);
    if (c->a.defer)
        launch(c, c->keep_fn == (const void*)&k_keep<2> ? (const void*)&k_keep_defer<2> : (const void*)&k_keep_defer<1>,
               c->a.keep_grid, q);
    else
        launch(c, c->keep_fn, c->grid, q);
}

// The grid of a deferred keep step: it waits on nothing, so it need not fit
// the resident slots as a look-back scan does, and a grid of at least
// keep_grid_min workgroups keeps its chunks at one sub-tile of kBlock
// positions up to that many sub-tiles (chunks of several sub-tiles loop
// through them in one workgroup).  The update launch that rebuilds the
// traversal plans the same chunks from KArgs::keep_grid.
uint32_t defer_keep_grid(const lsseg_ctx* c)
{
    return std::max<uint32_t>(c->grid, c->keep_grid_min);
}

// Before `count` more markings (issued by the caller after this, outside any
// graph capture): the markings write the stamp DevCtl::stamp into the
// emission records and every compaction advances it, so the stamps of the
// markings since the last reset are distinct as long as none passes 255.  If
// the next `count` would, every record is cleared and the stamp restarted so
// that the next marking writes 1; `pending`: the compaction of the last
// marking runs after this (in the next fused update launch) and advances the
// restarted stamp first.  c->stamp is the stamp of the next marking (each
// marking is followed by exactly one compaction); 0 after a device error.
void stamp_reserve(lsseg_ctx* c, uint32_t count, bool pending)
{
    if (c->stamp == 0 || c->stamp + count - 1u > 255u) {
        hipLaunchKernelGGL(k_records_reset, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a.C, c->a.g.N,
                           c->a.ctl, pending ? 0u : 1u);
        c->stamp = 1;
    }
    c->stamp += count;
}

// one complete iteration over list[q] (update, marking, keep, compaction into
// list[q ^ 1]); the context is at rest afterwards
void launch_iteration(lsseg_ctx* c, int q)
{
    stamp_reserve(c, 1, false);
    launch_head(c, q, false);
    launch_tail(c, q);
    launch(c, (const void*)&k_compact<kCompact>, c->grid, q);
    launch_check(c, q ^ 1);
}

void launch_update_only(lsseg_ctx* c, int op)
{
    void* args[] = {(void*)&c->a, (void*)&op};
    const void* fn = update_only_kernel(c->dtype, c->phi8);
    (void)hipLaunchKernel(fn, dim3(c->grid), dim3(kBlock), args, 0, c->stream);
}

// The padded geometry and the scan settings of a volume of X x Y x Z padded voxels.
void init_geom(Geom& g, uint32_t X, uint32_t Y, uint32_t Z)
{
    g.X = X;
    g.Y = Y;
    g.Z = Z;
    g.TX = g.X / 16;
    g.TY = g.Y / 16;
    g.TZ = g.Z / 16;
    g.sTY = g.TX * 4096u;
    g.sTZ = g.TX * g.TY * 4096u;
    g.N = (uint64_t)X * Y * Z;
    g.dTX = make_fastdiv(g.TX);
    g.dTXY = make_fastdiv(g.TX * g.TY);
    // few-path compactions of lists of up to 2^17 positions take tiles of
    // kBlock positions: four times the workgroups place the list, at the price
    // of a longer look-back (phantom 31.5 -> 29.6 ms; 2^16 and 2^17 measured
    // alike, 2^18 slower; tiles of 2 kBlock above gained nothing)
    g.small_n = (uint32_t)std::min<uint64_t>(g.N, (uint64_t)std::max(0, env_int("LSSEG_SMALL_TILE_N", 131072)));
    g.spec_tiles = (uint32_t)std::max(0, env_int("LSSEG_SPEC_TILES", (int)kSpecTiles));
    g.pack_counts = env_int("LSSEG_LOOKBACK_PACK", 1) ? 1u : 0u;
}

// look-back chunk states of a volume: one per keep or dense-compaction tile,
// seven per few-path compaction tile (group states: 7 per 64 of these, + 2)
uint64_t lookback_states(const Geom& g)
{
    return std::max<uint64_t>(g.N / std::min(kScanTile, kKeepTile) + 2,
                              7 * (std::max<uint64_t>(g.N / kPosTile, g.small_n / kBlock) + 2));
}

void fill_u64(lsseg_ctx* c, uint64_t* p, uint64_t count, uint64_t v)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(dense_grid(count)), dim3(kBlock), 0, c->stream, p, count, v);
}

// Dense compaction of the flags in C into list[target]; C stays non-zero
// exactly at the list's voxels, so afterwards the context is at rest.
lsseg_status dense_init_list(lsseg_ctx* c, int target)
{
    launch(c, (const void*)&k_compact<kInit>, c->grid, target ^ 1);
    if (env_int("LSSEG_CHECKS", 0))
        hipLaunchKernelGGL(k_check_list, dim3(c->sgrid), dim3(kBlock), 0, c->stream, c->a, target);
    HIP_TRY(c, hipGetLastError());
    c->q = target;
    return LSSEG_OK;
}

// `count` iterations starting from list parity q, at rest before and after.
// Each iteration's compaction runs in the next iteration's update launch
// (k_update_fused), the last one on its own: 3 count + 1 launches.  Runs of
// kGraphIters fused iterations (even, so the parity returns to its start)
// replay a captured graph.  LSSEG_FUSE=0 issues the four launches of every
// iteration separately instead (the same results).
lsseg_status launch_iterations(lsseg_ctx* c, int count, int q, uint32_t* skipped_out = nullptr)
{
    if (count <= 0) return LSSEG_OK;
    if (!env_int("LSSEG_FUSE", 1)) {
        for (int i = 0; i < count; ++i) launch_iteration(c, q ^ (i & 1));
        return LSSEG_OK;
    }
    // a deferred keep step leaves the traversal to the next update launch; the
    // chunks it counts follow the keep step's grid and width
    c->a.defer = c->defer_keep ? 1u : 0u;
    c->a.keep_grid = defer_keep_grid(c);
    c->a.keep_items = c->keep_fn == (const void*)&k_keep<2> ? 2u : 1u;
    struct DeferOff {
        lsseg_ctx* c;
        ~DeferOff() { c->a.defer = 0; }  // the other paths (records, re-initialisation) place their traversal
    } defer_off{c};
    stamp_reserve(c, 1, false);
    launch_head(c, q, false);
    launch_tail(c, q);
    q ^= 1;
    --count;
    const int chunk = env_int("LSSEG_GRAPH", 1) ? kGraphIters : 0;
    // An iteration over an empty list changes nothing but the counters, so once
    // a chunk ends with an empty list the remaining iterations are not issued:
    // k_empty_iterations writes their statistics and the parity follows from
    // the count.  The host reads the count of the chunk two before the one it
    // is about to issue, so the device always has a chunk queued meanwhile.
    int chunks = 0;
    uint32_t skipped = 0;
    while (chunk && !c->profile && count >= chunk) {
        const int slot = chunks & 1;
        if (chunks >= 2) {
            HIP_TRY(c, hipEventSynchronize(c->watch_ev[slot]));
            if (c->h_watch[slot] == 0) {
                skipped = (uint32_t)count;
                count = 0;
                break;
            }
        }
        hipGraphExec_t& ge = c->graphs[q];
        if (!ge) {
            hipGraph_t g = nullptr;
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < chunk; ++i) {
                launch_head(c, q ^ (i & 1), true);
                launch_tail(c, q ^ (i & 1));
            }
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ge = nullptr;
                return fail(c, LSSEG_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
            }
        }
        stamp_reserve(c, (uint32_t)chunk, true);
        HIP_TRY(c, hipGraphLaunch(ge, c->stream));
        HIP_TRY(c, hipMemcpyAsync(&c->h_watch[slot], &c->a.ctl->n[q], sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
        HIP_TRY(c, hipEventRecord(c->watch_ev[slot], c->stream));
        ++chunks;
        count -= chunk;
    }
    for (int i = 0; i < count; ++i) {
        stamp_reserve(c, 1, true);
        launch_head(c, q, true);
        launch_tail(c, q);
        q ^= 1;
        // a per-launch profiled run issues its launches one by one: it checks
        // the list count every kProfileWatch iterations and stops on an empty
        // list as the graph-replayed run does (the kernels' own start and stop
        // events are unaffected by the host wait)
        if (c->profile && i % kProfileWatch == kProfileWatch - 1 && i + 1 < count) {
            HIP_TRY(c, hipMemcpyAsync(&c->h_watch[0], &c->a.ctl->n[q], sizeof(uint32_t), hipMemcpyDeviceToHost,
                                      c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            if (c->h_watch[0] == 0) {
                skipped = (uint32_t)(count - i - 1);
                break;
            }
        }
    }
    launch(c, (const void*)&k_compact<kCompact>, c->grid, q ^ 1);
    if (c->a.defer)
        hipLaunchKernelGGL(k_settle_traversal, dim3(c->sgrid), dim3(kBlock), 0, c->stream, c->a, q);
    launch_check(c, q);
    if (skipped) hipLaunchKernelGGL(k_empty_iterations, dim3(1), dim3(kBlock), 0, c->stream, c->a, skipped);
    if (skipped_out) *skipped_out = skipped;
    return LSSEG_OK;
}

// text of a device error flag (kErr* bits)
const char* device_error_text(uint32_t err)
{
    return (err & kErrSpin)    ? "look-back wait exceeded its bound"
           : (err & kErrCheck) ? "list check failed: an entry outside the volume, repeated, or missing from the "
                                 "traversal"
           : (err & kErrCount) ? "the compaction's count differs from the keep step's"
           : (err & kErrStamp) ? "a marking found its emission stamp outside 1..255"
                               : "list overflow";
}

lsseg_status sync_and_check(lsseg_ctx* c)
{
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->h_ctl->error)
        return fail(c, LSSEG_ERR_DEVICE, "device error flag 0x%x (%s)", c->h_ctl->error,
                    device_error_text(c->h_ctl->error));
    return LSSEG_OK;
}

// Validates the parameters and derives the device constants for a source type
// (errors go to the context, or with c == nullptr to lsseg_last_error(NULL)).
lsseg_status make_params_for(lsseg_ctx* c, int32_t dtype, const lsseg_params* prm, DevParams* out);

lsseg_status make_params(lsseg_ctx* c, const lsseg_params* prm, DevParams* out)
{
    return make_params_for(c, c->dtype, prm, out);
}

lsseg_status make_params_for(lsseg_ctx* c, int32_t dtype, const lsseg_params* prm, DevParams* out)
{
    if (!prm) return fail(c, LSSEG_ERR_INVALID_ARG, "params is null");
    if (prm->mode != LSSEG_MODE_REF && prm->mode != LSSEG_MODE_CLEAN)
        return fail(c, LSSEG_ERR_INVALID_ARG, "unknown mode %d", prm->mode);
    if (!std::isfinite(prm->T) || !std::isfinite(prm->eps) || !std::isfinite(prm->alpha) ||
        !std::isfinite(prm->dt))
        return fail(c, LSSEG_ERR_INVALID_ARG, "non-finite parameter");
    DevParams P{};
    if (dtype == LSSEG_F32) {
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

lsseg_status ensure_stats(lsseg_ctx* c, uint64_t iters)
{
    if (iters <= c->stats_cap) return LSSEG_OK;
    const uint32_t cap = (uint32_t)std::max<uint64_t>(iters, 1024);
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
    return LSSEG_OK;
}

// Makes list[0] (n entries, already on the device) the current list: rebuilds
// the positions and the traversal S[0], taken from S_dev (device, n entries)
// if given, else the list sorted by tiled index.  The emission records need
// nothing: their bytes carry the stamps of earlier markings, which no later
// marking reuses before the records are cleared.
lsseg_status inject_list_counts(lsseg_ctx* c, uint32_t n, const uint32_t* S_dev = nullptr)
{
    c->q = 0;
    if (n) {
        hipLaunchKernelGGL(k_scatter_positions, dim3(dense_grid(n)), dim3(kBlock), 0, c->stream, c->a.list[0], n,
                           c->a.pos);
        if (S_dev) {
            HIP_TRY(c, hipMemcpyAsync(c->a.S[0], S_dev, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        } else {
            std::vector<uint32_t> h(n);
            HIP_TRY(c, hipMemcpyAsync(h.data(), c->a.list[0], n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            std::sort(h.begin(), h.end());
            HIP_TRY(c, hipMemcpyAsync(c->a.S[0], h.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
    }
    c->h_ctl->n[0] = n;
    c->h_ctl->n[1] = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl->n, c->h_ctl->n, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipGetLastError());
    return LSSEG_OK;
}

// a phi buffer as N floats (host memory); fixed point converts to k / 127
// The C ABI speaks the reference's tiled index tau; the engine stores voxel
// ids (Morton order inside each tile).  These convert list entries and whole
// padded volumes (host memory).
void taus_of_ids(uint32_t* v, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i) v[i] = tau_of_id(v[i]);
}

template <typename T>
void tau_order(T* vol, uint64_t N)  // id-ordered -> tau-ordered, in place
{
    std::vector<T> tmp(vol, vol + N);
    for (uint64_t t = 0; t < N; ++t) vol[t] = tmp[id_of_tau((uint32_t)t)];
}

template <typename T>
void id_order(T* vol, uint64_t N)  // tau-ordered -> id-ordered, in place
{
    std::vector<T> tmp(vol, vol + N);
    for (uint64_t t = 0; t < N; ++t) vol[id_of_tau((uint32_t)t)] = tmp[t];
}

lsseg_status phi_to_floats(lsseg_ctx* c, const void* buf, float* dst)
{
    const uint64_t N = c->a.g.N;
    if (!c->phi8) {
        HIP_TRY(c, hipMemcpy(dst, buf, N * sizeof(float), hipMemcpyDeviceToHost));
        return LSSEG_OK;
    }
    float* d = nullptr;
    lsseg_status st = dalloc(c, (void**)&d, N * sizeof(float));
    if (st != LSSEG_OK) return st;
    hipLaunchKernelGGL(k_phi8_to_float, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, (const int8_t*)buf, d, N);
    hipError_t e = hipMemcpyAsync(dst, d, N * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d);
    if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "phi readback failed: %s", hipGetErrorString(e));
    return LSSEG_OK;
}

// N floats (host memory) into a phi buffer; fixed point rounds v * 127
lsseg_status floats_to_phi(lsseg_ctx* c, const float* src, void* buf)
{
    const uint64_t N = c->a.g.N;
    if (!c->phi8) {
        HIP_TRY(c, hipMemcpy(buf, src, N * sizeof(float), hipMemcpyHostToDevice));
        return LSSEG_OK;
    }
    float* d = nullptr;
    lsseg_status st = dalloc(c, (void**)&d, N * sizeof(float));
    if (st != LSSEG_OK) return st;
    hipError_t e = hipMemcpyAsync(d, src, N * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        hipLaunchKernelGGL(k_float_to_phi8, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, (const float*)d,
                           (int8_t*)buf, N);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d);
    if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "phi upload failed: %s", hipGetErrorString(e));
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

}  // extern "C"

namespace {

void destroy_ctx(lsseg_ctx* ctx);

// lsseg_create on the given stream (nullptr: a stream of its own)
lsseg_status create_ctx(const lsseg_volume_desc* desc, const void* host_volume, hipStream_t shared,
                        lsseg_ctx** out_ctx)
{
    if (!desc || !host_volume || !out_ctx) return fail(nullptr, LSSEG_ERR_INVALID_ARG, "null argument");
    *out_ctx = nullptr;
    if (desc->nx == 0 || desc->ny == 0 || desc->nz == 0)
        return fail(nullptr, LSSEG_ERR_INVALID_ARG, "zero dimension");
    if (desc->dtype < LSSEG_U8 || desc->dtype > LSSEG_F32)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "unsupported dtype %d", desc->dtype);
    if (desc->phi_storage != LSSEG_PHI_F32 && desc->phi_storage != LSSEG_PHI_I8)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "unsupported phi storage %d", desc->phi_storage);
    const uint64_t X = (desc->nx + 15ull) / 16 * 16, Y = (desc->ny + 15ull) / 16 * 16,
                   Z = (desc->nz + 15ull) / 16 * 16;
    const uint64_t N = X * Y * Z;
    if (N >= (1ull << 32))
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "padded volume of %llu voxels exceeds the 32-bit index",
                    (unsigned long long)N);

    lsseg_ctx* c = new (std::nothrow) lsseg_ctx();
    if (!c) return fail(nullptr, LSSEG_ERR_OUT_OF_MEMORY, "host allocation failed");
    CreateAllocScope alloc_scope;
    auto bail = [&](lsseg_status st) {
        g_create_error = c->err;
        destroy_ctx(c);
        return st;
    };
    c->nx = desc->nx;
    c->ny = desc->ny;
    c->nz = desc->nz;
    c->dtype = desc->dtype;
    c->S = (desc->dtype <= LSSEG_I8) ? 1 : (desc->dtype <= LSSEG_I16 ? 2 : 4);
    c->z_aniso = desc->z_aniso;
    c->phi8 = desc->phi_storage == LSSEG_PHI_I8;
    c->P = c->phi8 ? 1 : sizeof(float);
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        c->err = std::string("no HIP device: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    c->cus = (uint32_t)std::max(1, cus);
    // the keep and compaction kernels (and the update fused with a compaction)
    // traversal positions per thread of the keep step: one, or two with
    // LSSEG_KEEP_ITEMS=2 (one measured 4% faster on the 181x217x181 phantom)
    c->keep_fn = env_int("LSSEG_KEEP_ITEMS", 1) == 2 ? (const void*)&k_keep<2> : (const void*)&k_keep<1>;
    const int occ = min_occupancy({(const void*)&k_compact<kCompact>, (const void*)&k_compact<kInit>, c->keep_fn,
                                   update_fused_kernel(desc->dtype, c->phi8)});
    const int want = env_int("LSSEG_GRID_BLOCKS", 0);
    c->grid = want > 0 ? (uint32_t)want : (uint32_t)occ * c->cus;
    c->grid = std::max<uint32_t>(16u, c->grid & ~7u);  // k_update_fused splits it in multiples of 8
    c->grid0 = std::max<uint32_t>(16u, ((uint32_t)occ * c->cus) & ~7u);
    // k_update_fused: at most three quarters of its grid compact, the rest
    // updates (by default its grid is the keep/compaction grid)
    const int fwant = env_int("LSSEG_FUSED_GRID", 0);
    c->fgrid = fwant > 0 ? std::max<uint32_t>(16u, ((uint32_t)fwant + 7u) & ~7u) : c->grid;
    c->a.fused_compact_max = (c->fgrid / 4u * 3u) & ~7u;
    // update and marking: eight workgroups per CU (a multiple of the eight
    // XCDs), list positions dealt to the XCDs in runs of 4096
    const int swant = env_int("LSSEG_STREAM_BLOCKS", 0);
    c->sgrid = swant > 0 ? ((uint32_t)swant + 7u) / 8u * 8u : 8u * c->cus;
    const int run = env_int("LSSEG_XCD_RUN", 4096);
    c->a.run_log2 = 0;
    if (run >= kBlock) {
        uint32_t lr = 0;
        while ((1u << (lr + 1)) <= (uint32_t)run && lr < 24) ++lr;
        c->a.run_log2 = lr;
    }
    if (shared) {
        c->stream = shared;
        c->owns_stream = false;
    } else if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        c->err = std::string("stream creation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        c->err = std::string("stream/event creation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    if (hipHostMalloc((void**)&c->h_ctl, sizeof(DevCtl), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->h_watch, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        c->err = "pinned allocation failed";
        return bail(LSSEG_ERR_OUT_OF_MEMORY);
    }
    for (hipEvent_t& ev : c->watch_ev)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            c->err = std::string("event creation failed: ") + hipGetErrorString(e);
            return bail(LSSEG_ERR_DEVICE);
        }
    memset(c->h_ctl, 0, sizeof(DevCtl));

    Geom& g = c->a.g;
    init_geom(g, (uint32_t)X, (uint32_t)Y, (uint32_t)Z);
    c->small_n_max = g.small_n;
    c->defer_keep = !shared && env_int("LSSEG_DEFER_KEEP", 1) != 0;
    c->keep_grid_min = (uint32_t)std::max(16, env_int("LSSEG_KEEP_GRID", 4096));

    void* p = nullptr;
    const uint64_t ntiles = lookback_states(g);
    lsseg_status st;
#define ALLOC(field, type, count)                                          \
    if ((st = dalloc(c, &p, (size_t)(count) * sizeof(type))) != LSSEG_OK) \
        return bail(st);                                                   \
    c->a.field = (type*)p;
    ALLOC(phi[0], uint8_t, N * c->P);
    ALLOC(phi[1], uint8_t, N * c->P);
    ALLOC(moved, uint8_t, N);
    ALLOC(E, int8_t, N);
    ALLOC(frozen, int8_t, N);
    ALLOC(C, uint64_t, N);
    ALLOC(list[0], uint32_t, N);
    ALLOC(list[1], uint32_t, N);
    ALLOC(S[0], uint32_t, N);
    ALLOC(S[1], uint32_t, N);
    ALLOC(pos, uint32_t, N);
    ALLOC(emit, uint8_t, N);
    ALLOC(keep, uint8_t, N);
    ALLOC(memb, uint8_t, N);
    ALLOC(tstate, uint64_t, ntiles);
    ALLOC(gstate, uint64_t, 7 * (ntiles / 64 + 2));
    ALLOC(cpart, uint64_t, kCountSlots * kCountStride);
    ALLOC(dirc, uint32_t, kDirCounters);
    ALLOC(kcnt, uint32_t, N / kBlock + 64);  // a keep chunk holds at least one tile of kBlock positions
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
    if (device_malloc(&lin, lin_bytes) != hipSuccess) {
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
    fill_phi_one(c, c->a.phi[0]);
    fill_phi_one(c, c->a.phi[1]);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.E, N, (uint8_t)127);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.frozen, N, (uint8_t)127);
    fill_u64(c, c->a.C, N, 0);
    fill_u64(c, c->a.tstate, ntiles, 0);
    fill_u64(c, c->a.gstate, 7 * (ntiles / 64 + 2), 0);
    fill_u64(c, c->a.cpart, kCountSlots * kCountStride, 0);
    fill_u64(c, (uint64_t*)c->a.dirc, kDirCounters / 2, 0);
    // positions, masks and keep bytes of voxels outside the list are never read;
    // zero keeps every entry a valid index
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.pos, N * sizeof(uint32_t),
                       (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.emit, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.keep, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, c->a.memb, N, (uint8_t)0);
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

void destroy_ctx(lsseg_ctx* ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    detach_records(ctx);
    destroy_graphs(ctx);
    for (hipEvent_t e : ctx->prof_events) (void)hipEventDestroy(e);
    ctx->prof_events.clear();
    for (void* p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
    if (ctx->h_watch) (void)hipHostFree(ctx->h_watch);
    for (hipEvent_t ev : ctx->watch_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream && ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // namespace

extern "C" {

lsseg_status lsseg_create(const lsseg_volume_desc* desc, const void* host_volume, lsseg_ctx** out_ctx)
{
    return create_ctx(desc, host_volume, nullptr, out_ctx);
}

lsseg_status lsseg_destroy(lsseg_ctx* ctx)
{
    if (!ctx) return LSSEG_ERR_INVALID_ARG;
    if (ctx->batch) return fail(ctx, LSSEG_ERR_STATE, "this context belongs to a batch; destroy the batch");
    destroy_ctx(ctx);
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
    hipError_t e = hipMemcpyAsync(lin, phi0, cnt * sizeof(float), hipMemcpyDefault, c->stream);
    if (e == hipSuccess) {
        e = launch_phi(c, &k_set_phi<float>, &k_set_phi<int8_t>, dense_grid(c->a.g.N), c->a, (const float*)lin,
                       c->nx, c->ny, c->nz, band);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(lin);
    if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "set_phi failed: %s", hipGetErrorString(e));
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
        HIP_TRY(c, launch_phi(c, &k_seed_pass<float>, &k_seed_pass<int8_t>, dg, c->a, rd, sx, sy, sz,
                              (int32_t)radius, c->z_aniso));
        rd ^= 1;
    }
    HIP_TRY(c, launch_phi(c, &k_keep_flags<float>, &k_keep_flags<int8_t>, dg, c->a, rd));
    HIP_TRY(c, hipGetLastError());
    // two swaps leave the read parity unchanged: the list goes to that parity
    lsseg_status st = dense_init_list(c, rd);
    if (st != LSSEG_OK) return st;
    return sync_and_check(c);
}

lsseg_status lsseg_iterate_reinit(lsseg_ctx* c, const lsseg_params* prm)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    DevParams P;
    lsseg_status st = make_params(c, prm, &P);
    if (st != LSSEG_OK) return st;
    if ((st = upload_params(c, P)) != LSSEG_OK) return st;
    launch_update_only(c, c->q);
    const int rd = c->q ^ 1;
    HIP_TRY(c, launch_phi(c, &k_keep_flags<float>, &k_keep_flags<int8_t>, dense_grid(c->a.g.N), c->a, rd));
    HIP_TRY(c, hipGetLastError());
    if ((st = dense_init_list(c, rd)) != LSSEG_OK) return st;
    return sync_and_check(c);
}

lsseg_status lsseg_iterate(lsseg_ctx* c, const lsseg_params* prm, int32_t iters, lsseg_stats* stats)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    if (iters < 0) return fail(c, LSSEG_ERR_INVALID_ARG, "negative iteration count");
    DevParams P;
    lsseg_status st = make_params(c, prm, &P);
    if (st != LSSEG_OK) return st;
    const bool checks = env_int("LSSEG_CHECKS", 0) != 0;
    if (checks != c->checks) {
        c->checks = checks;
        destroy_graphs(c);  // captured without (or with) the check launches
    }
    if ((st = ensure_stats(c, (uint64_t)iters)) != LSSEG_OK) return st;
    if ((st = upload_params(c, P)) != LSSEG_OK) return st;
    // reset the per-call counters (keep n[], epoch)
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->h_ctl->iter = 0;
    c->h_ctl->stats_cap = c->stats_cap;
    // after a call that ended with a device error, look-back states of its
    // unfinished scans may carry the current epoch (skip past it), and an
    // unfinished compaction may have left keep bytes set (clear them)
    if (c->h_ctl->error) {
        c->h_ctl->epoch += 2;
        hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a.keep, c->a.g.N,
                           (uint8_t)0);
        c->stamp = 0;  // the records are cleared before the next marking
    }
    c->h_ctl->error = 0;
    c->h_ctl->few_iters = 0;
    c->h_ctl->many_iters = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl, c->h_ctl, sizeof(DevCtl), hipMemcpyHostToDevice, c->stream));

    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    uint32_t skipped = 0;
    if (iters > 0) {
        const int q0 = c->q;
        if ((st = launch_iterations(c, iters, q0, &skipped)) != LSSEG_OK) return st;
        c->q = q0 ^ (iters & 1);
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    if ((st = sync_and_check(c)) != LSSEG_OK) return st;
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
        stats->empty_skipped = skipped;
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

lsseg_status lsseg_iterate_profiled(lsseg_ctx* c, const lsseg_params* prm, int32_t iters, lsseg_stats* stats,
                                    lsseg_kernel_times* times, float* launch_ms, uint64_t launch_cap)
{
    if (!c || !times) return LSSEG_ERR_INVALID_ARG;
    if (iters < 0) return fail(c, LSSEG_ERR_INVALID_ARG, "negative iteration count");
    const size_t need = 8 * (size_t)iters + 2;
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
    if (iters == 0) return LSSEG_OK;
    // a start/stop event pair per launch, in issue order: the update, marking
    // and keep step of each iteration (the update from the second iteration on
    // fused with the previous compaction), then the last compaction
    const int32_t launches = (int32_t)(c->prof_next / 2);
    for (int32_t i = 0; i < launches; ++i) {
        float ms = 0;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->prof_events[2 * i], c->prof_events[2 * i + 1]));
        times->total_ms += ms;
        if (i == 0) times->first_ms = ms;
        if (i == launches - 1) times->last_ms = ms;
        if (launch_ms && (uint64_t)i < launch_cap) launch_ms[i] = ms;
    }
    times->launches = (uint64_t)launches;
    return LSSEG_OK;
}

lsseg_status lsseg_get_iteration_counts(const lsseg_ctx* c, uint32_t* n, uint64_t* cc, uint32_t* m, uint64_t cap,
                                        uint64_t* count)
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
    uint32_t* v = &cm->h_ctl->n[c->q];
    HIP_TRY(cm, hipMemcpyAsync(v, &c->a.ctl->n[c->q], sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(cm, hipStreamSynchronize(c->stream));
    *n = *v;
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
    taus_of_ids(out, k);
    return cap < cnt ? fail(cm, LSSEG_ERR_CAPACITY, "list has %llu entries, capacity %llu", (unsigned long long)cnt,
                            (unsigned long long)cap)
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
    hipError_t e = launch_phi(cm, &k_get_phi<float>, &k_get_phi<int8_t>, dense_grid(cnt), c->a,
                              which == 0 ? c->q : c->q ^ 1, d, c->nx, c->ny, c->nz);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, cnt * sizeof(float), hipMemcpyDefault, c->stream);
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
    hipLaunchKernelGGL(k_get_mask, dim3(dense_grid(cnt)), dim3(kBlock), 0, c->stream, c->a, d, c->nx, c->ny, c->nz);
    hipError_t e = hipMemcpyAsync(out, d, cnt, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(cm, LSSEG_ERR_DEVICE, "get_mask failed: %s", hipGetErrorString(e));
    return LSSEG_OK;
}

lsseg_status lsseg_auto_parameters(lsseg_ctx* c, float* T, float* eps, uint64_t* count)
{
    if (!c || !T || !eps) return LSSEG_ERR_INVALID_ARG;
    const void* fn;
    switch (c->dtype) {
    case LSSEG_U8: fn = (const void*)&k_inside_moments<uint8_t, false>; break;
    case LSSEG_I8: fn = (const void*)&k_inside_moments<int8_t, false>; break;
    case LSSEG_U16: fn = (const void*)&k_inside_moments<uint16_t, false>; break;
    case LSSEG_I16: fn = (const void*)&k_inside_moments<int16_t, false>; break;
    case LSSEG_U32: fn = (const void*)&k_inside_moments<uint32_t, false>; break;
    case LSSEG_I32: fn = (const void*)&k_inside_moments<int32_t, false>; break;
    default: fn = (const void*)&k_inside_moments<float, true>; break;
    }
    const uint32_t dg = dense_grid(c->a.g.N);
    void* base = nullptr;
    lsseg_status st;
    if ((st = dalloc(c, &base, sizeof(unsigned long long) + (size_t)dg * sizeof(float))) != LSSEG_OK) return st;
    unsigned long long* d_cnt = (unsigned long long*)base;
    float* d_part = (float*)((char*)base + sizeof(unsigned long long));  // one partial sum per workgroup
    std::vector<float> part(dg);
    unsigned long long cnt = 0;
    hipError_t e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), c->stream);
    // as the reference: FP32 sums, the mean as sum / (float)count, then the
    // squared deviations from it
    float mean = 0.0f, sums[2] = {0.0f, 0.0f};
    for (int mode = 0; mode < 2 && e == hipSuccess; ++mode) {
        void* args[] = {(void*)&c->a, (void*)&mode, (void*)&mean, (void*)&d_part, (void*)&d_cnt};
        e = hipLaunchKernel(fn, dim3(dg), dim3(kBlock), args, 0, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(part.data(), d_part, (size_t)dg * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        sums[mode] = pairwise_sum(part.data(), dg);
        if (mode == 0 && cnt) mean = sums[0] / (float)cnt;
    }
    dfree(c, base);
    if (e != hipSuccess) return fail(c, LSSEG_ERR_DEVICE, "auto parameters failed: %s", hipGetErrorString(e));
    if (count) *count = cnt;
    if (cnt < 2) return fail(c, LSSEG_ERR_STATE, "fewer than two voxels inside the segmentation");
    *T = mean;
    *eps = std::sqrt(sums[1] / ((float)cnt - 1.0f));
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
    fill_phi_one(c, c->a.phi[0]);
    fill_phi_one(c, c->a.phi[1]);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.E, N, (uint8_t)127);
    if (all)
        hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, c->stream, (uint8_t*)c->a.frozen, N, (uint8_t)127);
    lsseg_status st = inject_list_counts(c, 0);
    if (st != LSSEG_OK) return st;
    return sync_and_check(c);
}

lsseg_status lsseg_get_state(const lsseg_ctx* c, float* phi_read, float* phi_write, int8_t* E, uint32_t* list,
                             uint64_t list_cap, uint64_t* n)
{
    if (!c) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* cm = const_cast<lsseg_ctx*>(c);
    const uint64_t N = c->a.g.N;
    for (int i = 0; i < 2; ++i) {
        float* dst = i == 0 ? phi_read : phi_write;
        if (!dst) continue;
        lsseg_status st = phi_to_floats(cm, c->a.phi[c->q ^ i], dst);
        if (st != LSSEG_OK) return st;
        tau_order(dst, N);
    }
    if (E) {
        HIP_TRY(cm, hipMemcpy(E, c->a.E, N, hipMemcpyDeviceToHost));
        tau_order(E, N);
    }
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
    std::vector<uint32_t> ids(n);
    for (uint64_t i = 0; i < n; ++i) {
        if (list[i] >= N) return fail(c, LSSEG_ERR_INVALID_ARG, "list entry %llu out of range", (unsigned long long)i);
        if (seen[list[i]])
            return fail(c, LSSEG_ERR_INVALID_ARG, "list entry %llu is a duplicate", (unsigned long long)i);
        seen[list[i]] = 1;
        ids[i] = id_of_tau(list[i]);
    }
    std::vector<float> ph(phi_read, phi_read + N);
    id_order(ph.data(), N);
    lsseg_status st0 = floats_to_phi(c, ph.data(), c->a.phi[0]);
    if (st0 != LSSEG_OK) return st0;
    ph.assign(phi_write, phi_write + N);
    id_order(ph.data(), N);
    if ((st0 = floats_to_phi(c, ph.data(), c->a.phi[1])) != LSSEG_OK) return st0;
    std::vector<int8_t> eh(E, E + N);
    id_order(eh.data(), N);
    HIP_TRY(c, hipMemcpy(c->a.E, eh.data(), N, hipMemcpyHostToDevice));
    if (n) HIP_TRY(c, hipMemcpy(c->a.list[0], ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    lsseg_status st = inject_list_counts(c, (uint32_t)n);
    if (st != LSSEG_OK) return st;
    return sync_and_check(c);
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
        if ((st = dalloc(c, &p, N * c->P)) != LSSEG_OK) return st;
        k.phi[0] = p;
        if ((st = dalloc(c, &p, N * c->P)) != LSSEG_OK) return st;
        k.phi[1] = p;
        if ((st = dalloc(c, &p, N)) != LSSEG_OK) return st;
        k.E = (int8_t*)p;
        if ((st = dalloc(c, &p, N * sizeof(uint32_t))) != LSSEG_OK) return st;
        k.list = (uint32_t*)p;
        if ((st = dalloc(c, &p, N * sizeof(uint32_t))) != LSSEG_OK) return st;
        k.S = (uint32_t*)p;
    }
    uint64_t n = 0;
    lsseg_status st = lsseg_get_active_count(c, &n);
    if (st != LSSEG_OK) return st;
    HIP_TRY(c, hipMemcpyAsync(k.phi[0], c->a.phi[c->q], N * c->P, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(k.phi[1], c->a.phi[c->q ^ 1], N * c->P, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(k.E, c->a.E, N, hipMemcpyDeviceToDevice, c->stream));
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(k.list, c->a.list[c->q], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(k.S, c->a.S[c->q], n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    }
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
    HIP_TRY(c, hipMemcpyAsync(c->a.phi[0], k.phi[0], N * c->P, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->a.phi[1], k.phi[1], N * c->P, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->a.E, k.E, N, hipMemcpyDeviceToDevice, c->stream));
    if (k.n) HIP_TRY(c, hipMemcpyAsync(c->a.list[0], k.list, k.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    lsseg_status st = inject_list_counts(c, k.n, k.S);
    if (st != LSSEG_OK) return st;
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

lsseg_status lsseg_probe_random_access(void* buf, uint64_t bytes, uint32_t count, int32_t mode, int32_t reps,
                                       void* stream, float* us_median)
{
    if (!buf || !us_median || count == 0 || reps < 1 || reps > 1000 || bytes < 256 || (bytes & 3u) ||
        (reinterpret_cast<uintptr_t>(buf) & 3u) || (mode != 0 && mode != 1 && mode != 2))
        return fail(nullptr, LSSEG_ERR_INVALID_ARG, "lsseg_probe_random_access: invalid argument");
    const void* fn = mode == 0   ? (const void*)&k_probe_access<1, false>
                     : mode == 1 ? (const void*)&k_probe_access<7, false>
                                 : (const void*)&k_probe_access<1, true>;
    hipStream_t s = (hipStream_t)stream;
    float* a = static_cast<float*>(buf);
    float* sink = a;  // written only where a thread's loads sum to the kernel's sentinel (never on a zeroed buffer)
    uint64_t words = bytes / 4u;
    std::vector<float> ms;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int32_t r = 0; r < reps && e == hipSuccess; ++r) {
        uint32_t seed = 0x9e3779b9u * (uint32_t)(r + 1);
        void* args[] = {(void*)&a, (void*)&words, (void*)&count, (void*)&seed, (void*)&sink};
        e = hipExtLaunchKernel(fn, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), args, 0, s, e0, e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float m = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&m, e0, e1);
        ms.push_back(m);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(nullptr, LSSEG_ERR_DEVICE, "lsseg_probe_random_access: %s", hipGetErrorString(e));
    std::sort(ms.begin(), ms.end());
    *us_median = 1e3f * ms[ms.size() / 2];
    return LSSEG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// one segmentation run that allocates nothing (lsseg_segment_noalloc)
// ---------------------------------------------------------------------------
// Every buffer of the run except the caller's inputs and outputs lives in one
// static device array per GPU.  Its address is looked up once per device with
// hipGetSymbolAddress (host code never passes the array's name to a kernel or
// a copy: on the host that name is only a stand-in), and the run is carved
// from that address.
namespace {

constexpr uint64_t kNoAllocVoxels = LSSEG_NOALLOC_MAX_VOXELS;
constexpr uint64_t kNoAllocIters = LSSEG_NOALLOC_MAX_ITERS;

// buffer offsets of a run of N padded voxels (256-byte aligned)
struct NoAllocLayout {
    uint64_t phi[2], moved, E, frozen, C, list[2], S[2], pos, emit, keep, memb, src;
    uint64_t tstate, gstate, cpart, dirc, ctl, prm, sums, n_hist, c_hist, m_hist;
    uint64_t end;
};

constexpr uint64_t align_up(uint64_t v) { return (v + 255u) & ~uint64_t(255); }

NoAllocLayout noalloc_layout(const Geom& g)
{
    const uint64_t N = g.N;
    const uint64_t ntiles = lookback_states(g);
    NoAllocLayout L{};
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = align_up(at + bytes);
        return o;
    };
    L.phi[0] = take(N * 4);
    L.phi[1] = take(N * 4);
    L.moved = take(N);
    L.E = take(N);
    L.frozen = take(N);
    L.C = take(N * 8);
    L.list[0] = take(N * 4);
    L.list[1] = take(N * 4);
    L.S[0] = take(N * 4);
    L.S[1] = take(N * 4);
    L.pos = take(N * 4);
    L.emit = take(N);
    L.keep = take(N);
    L.memb = take(N);
    L.src = take(N * 4);
    L.tstate = take(ntiles * 8);
    L.gstate = take(7 * (ntiles / 64 + 2) * 8);
    L.cpart = take((uint64_t)kCountSlots * kCountStride * 8);
    L.dirc = take(kDirCounters * 4);
    L.ctl = take(sizeof(DevCtl));
    L.prm = take(sizeof(DevParams));
    L.sums = take(sizeof(IterSums));
    L.n_hist = take(kNoAllocIters * 4);
    L.c_hist = take(kNoAllocIters * 8);
    L.m_hist = take(kNoAllocIters * 4);
    L.end = at;
    return L;
}

// the largest run's layout: a volume of kNoAllocVoxels voxels whose padded
// dims are 16 x 16 x (kNoAllocVoxels / 256) has the most look-back states
constexpr uint64_t kNoAllocBytes = 47ull * kNoAllocVoxels + (2ull << 20);
static_assert(LSSEG_NOALLOC_SCRATCH_BYTES == kNoAllocBytes, "header and library agree on the scratch size");

}  // namespace

// the scratch array: one per GPU, held until the process ends
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
    uint8_t* ptr = nullptr;  // device address of `scratch` on this GPU
    uint32_t grid = 0, sgrid = 0;
};
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocDevices];

// the scratch address and launch grids of the current device, set up on the
// first call on it
lsseg_status noalloc_device(NoAllocDevice* out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, LSSEG_ERR_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocDevices) return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, LSSEG_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cus = std::max(1, cus);
        const int occ = min_occupancy({(const void*)&k_compact<kCompact>, (const void*)&k_compact<kInit>,
                                       (const void*)&k_keep<1>, update_fused_kernel(LSSEG_F32, false),
                                       update_fused_kernel(LSSEG_U8, false), update_fused_kernel(LSSEG_U16, false)});
        d.grid = std::max<uint32_t>(16u, ((uint32_t)occ * (uint32_t)cus) & ~7u);
        d.sgrid = 8u * (uint32_t)cus;
        d.ptr = (uint8_t*)p;
    }
    *out = d;
    return LSSEG_OK;
}

}  // namespace

extern "C" {

lsseg_status lsseg_segment_noalloc(const lsseg_volume_desc* desc, const void* src, const float* phi0, float band,
                                   const lsseg_params* params, int32_t iters, float* phi_out, uint8_t* mask_out,
                                   void* stream, lsseg_stats* stats)
{
    if (!desc || !src || !phi0 || !params || !phi_out || !mask_out)
        return fail(nullptr, LSSEG_ERR_INVALID_ARG, "null argument");
    if (desc->nx == 0 || desc->ny == 0 || desc->nz == 0) return fail(nullptr, LSSEG_ERR_INVALID_ARG, "zero dimension");
    if (desc->dtype < LSSEG_U8 || desc->dtype > LSSEG_F32)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "unsupported dtype %d", desc->dtype);
    if (desc->phi_storage != LSSEG_PHI_F32 && desc->phi_storage != LSSEG_PHI_I8)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "unsupported phi storage %d", desc->phi_storage);
    if (!(band > 0.0f)) return fail(nullptr, LSSEG_ERR_INVALID_ARG, "band must be positive");
    if (iters < 0 || (uint64_t)iters > kNoAllocIters)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "lsseg_segment_noalloc runs 0..%llu iterations, not %d",
                    (unsigned long long)kNoAllocIters, iters);
    const uint64_t X = (desc->nx + 15ull) / 16 * 16, Y = (desc->ny + 15ull) / 16 * 16,
                   Z = (desc->nz + 15ull) / 16 * 16;
    if (X * Y * Z > kNoAllocVoxels)
        return fail(nullptr, LSSEG_ERR_UNSUPPORTED,
                    "lsseg_segment_noalloc takes at most %llu padded voxels (its static scratch holds %llu bytes); "
                    "%llux%llux%llu pads to %llu: use lsseg_create",
                    (unsigned long long)kNoAllocVoxels, (unsigned long long)kNoAllocBytes,
                    (unsigned long long)desc->nx, (unsigned long long)desc->ny, (unsigned long long)desc->nz,
                    (unsigned long long)(X * Y * Z));
    DevParams P;
    lsseg_status st = make_params_for(nullptr, desc->dtype, params, &P);
    if (st != LSSEG_OK) return st;
    NoAllocDevice dev;
    if ((st = noalloc_device(&dev)) != LSSEG_OK) return st;

    KArgs a{};
    init_geom(a.g, (uint32_t)X, (uint32_t)Y, (uint32_t)Z);
    const NoAllocLayout L = noalloc_layout(a.g);
    if (L.end > kNoAllocBytes) return fail(nullptr, LSSEG_ERR_UNSUPPORTED, "scratch layout exceeds the array");
    uint8_t* base = dev.ptr;
    a.run_log2 = 12;  // list positions dealt to the XCDs in runs of 4096, as lsseg_create does
    a.fused_compact_max = (dev.grid / 4u * 3u) & ~7u;
    a.phi[0] = base + L.phi[0];
    a.phi[1] = base + L.phi[1];
    a.moved = base + L.moved;
    a.src = base + L.src;
    a.E = (int8_t*)(base + L.E);
    a.frozen = (int8_t*)(base + L.frozen);
    a.C = (uint64_t*)(base + L.C);
    for (int i = 0; i < 2; ++i) {
        a.list[i] = (uint32_t*)(base + L.list[i]);
        a.S[i] = (uint32_t*)(base + L.S[i]);
    }
    a.pos = (uint32_t*)(base + L.pos);
    a.emit = base + L.emit;
    a.keep = base + L.keep;
    a.memb = base + L.memb;
    a.tstate = (uint64_t*)(base + L.tstate);
    a.gstate = (uint64_t*)(base + L.gstate);
    a.cpart = (uint64_t*)(base + L.cpart);
    a.dirc = (uint32_t*)(base + L.dirc);
    a.ctl = (DevCtl*)(base + L.ctl);
    a.prm = (const DevParams*)(base + L.prm);
    a.n_hist = (uint32_t*)(base + L.n_hist);
    a.c_hist = (uint64_t*)(base + L.c_hist);
    a.m_hist = (uint32_t*)(base + L.m_hist);

    hipStream_t s = (hipStream_t)stream;
    const bool phi8 = desc->phi_storage == LSSEG_PHI_I8;
    const uint32_t S = (desc->dtype <= LSSEG_I8) ? 1 : (desc->dtype <= LSSEG_I16 ? 2 : 4);
    const bool is_signed = desc->dtype == LSSEG_I8 || desc->dtype == LSSEG_I16 || desc->dtype == LSSEG_I32;
    const uint64_t N = a.g.N;
    const uint32_t dg = dense_grid(N);
    const uint64_t ntiles = lookback_states(a.g);
    // the scratch keeps the previous call's contents: everything a run reads
    // before writing it is set here, with kernels only (no memset nodes, so
    // the calls can be captured in a graph)
    hipLaunchKernelGGL(k_pad_tile_source, dim3(dg), dim3(kBlock), 0, s, a.g, (const uint8_t*)src, (uint8_t*)a.src,
                       desc->nx, desc->ny, desc->nz, S, (uint8_t)(is_signed ? 0x81 : 0x00));
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, s, (uint8_t*)a.frozen, N, (uint8_t)127);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, s, (uint8_t*)a.pos, N * sizeof(uint32_t), (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, s, a.emit, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, s, a.keep, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u8, dim3(dg), dim3(kBlock), 0, s, a.memb, N, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u64, dim3(dense_grid(ntiles)), dim3(kBlock), 0, s, a.tstate, ntiles, (uint64_t)0);
    hipLaunchKernelGGL(k_fill_u64, dim3(dense_grid(ntiles)), dim3(kBlock), 0, s, a.gstate, 7 * (ntiles / 64 + 2),
                       (uint64_t)0);
    hipLaunchKernelGGL(k_fill_u64, dim3(4), dim3(kBlock), 0, s, a.cpart, (uint64_t)kCountSlots * kCountStride,
                       (uint64_t)0);
    hipLaunchKernelGGL(k_fill_u64, dim3(dense_grid(kDirCounters / 2)), dim3(kBlock), 0, s, (uint64_t*)a.dirc,
                       (uint64_t)kDirCounters / 2, (uint64_t)0);
    DevCtl c0{};
    c0.stats_cap = (uint32_t)iters;
    c0.stamp = 1;  // set_phi leaves flags only at the list's voxels, which the first update clears
    hipLaunchKernelGGL(k_put_params_ctl, dim3(1), dim3(64), 0, s, (DevParams*)a.prm, P, a.ctl, c0);
    // lsseg_set_phi: both phi buffers, export byte, band flags, then the
    // dense compaction of the flags into list[0]
    if (phi8)
        hipLaunchKernelGGL(k_set_phi<int8_t>, dim3(dg), dim3(kBlock), 0, s, a, phi0, desc->nx, desc->ny, desc->nz,
                           band);
    else
        hipLaunchKernelGGL(k_set_phi<float>, dim3(dg), dim3(kBlock), 0, s, a, phi0, desc->nx, desc->ny, desc->nz,
                           band);
    {
        int q1 = 1;
        void* args[] = {(void*)&a, (void*)&q1};
        (void)hipLaunchKernel((const void*)&k_compact<kInit>, dim3(dev.grid), dim3(kBlock), args, 0, s);
    }
    // lsseg_iterate: the update, marking and keep step of each iteration, each
    // compaction in the next update's launch, the last one on its own
    int q = 0;
    uint32_t stamp = 1;  // as stamp_reserve
    for (int32_t i = 0; i < iters; ++i) {
        if (stamp > 255u) {  // iteration i > 0: the last compaction is in this iteration's update launch
            hipLaunchKernelGGL(k_records_reset, dim3(dg), dim3(kBlock), 0, s, a.C, N, a.ctl, 0u);
            stamp = 1;
        }
        ++stamp;
        void* args[] = {(void*)&a, (void*)&q};
        const void* head = i ? update_fused_kernel(desc->dtype, phi8) : update_kernel(desc->dtype, phi8);
        (void)hipLaunchKernel(head, dim3(i ? dev.grid : dev.sgrid), dim3(kBlock), args, 0, s);
        (void)hipLaunchKernel(mark_kernel(phi8), dim3(dev.sgrid), dim3(kBlock), args, 0, s);
        (void)hipLaunchKernel((const void*)&k_keep<1>, dim3(dev.grid), dim3(kBlock), args, 0, s);
        q ^= 1;
    }
    if (iters > 0) {
        int qc = q ^ 1;
        void* args[] = {(void*)&a, (void*)&qc};
        (void)hipLaunchKernel((const void*)&k_compact<kCompact>, dim3(dev.grid), dim3(kBlock), args, 0, s);
    }
    // lsseg_get_phi(0) and lsseg_get_mask into the caller's buffers
    const uint64_t cnt = (uint64_t)desc->nx * desc->ny * desc->nz;
    if (phi8)
        hipLaunchKernelGGL(k_get_phi<int8_t>, dim3(dense_grid(cnt)), dim3(kBlock), 0, s, a, q, phi_out, desc->nx,
                           desc->ny, desc->nz);
    else
        hipLaunchKernelGGL(k_get_phi<float>, dim3(dense_grid(cnt)), dim3(kBlock), 0, s, a, q, phi_out, desc->nx,
                           desc->ny, desc->nz);
    hipLaunchKernelGGL(k_get_mask, dim3(dense_grid(cnt)), dim3(kBlock), 0, s, a, mask_out, desc->nx, desc->ny,
                       desc->nz);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, LSSEG_ERR_DEVICE, "launch failed: %s", hipGetErrorString(e));
    if (!stats) return LSSEG_OK;
    // statistics: summed on the device, then one copy and a wait
    IterSums* d_sums = (IterSums*)(base + L.sums);
    hipLaunchKernelGGL(k_sum_history, dim3(1), dim3(kBlock), 0, s, a.n_hist, a.c_hist, a.m_hist, (uint32_t)iters,
                       d_sums);
    struct {
        DevCtl ctl;
        IterSums sums;
    } h{};
    e = hipMemcpyAsync(&h.ctl, a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&h.sums, d_sums, sizeof(IterSums), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(nullptr, LSSEG_ERR_DEVICE, "statistics readback failed: %s", hipGetErrorString(e));
    if (h.ctl.error)
        return fail(nullptr, LSSEG_ERR_DEVICE, "device error flag 0x%x (%s)", h.ctl.error, device_error_text(h.ctl.error));
    memset(stats, 0, sizeof(*stats));
    stats->iterations = (uint64_t)iters;
    stats->sum_active = h.sums.n;
    stats->sum_candidates = h.sums.c;
    stats->sum_new_active = h.sums.m;
    stats->active_final = h.ctl.n[q];
    stats->few_path_iters = h.ctl.few_iters;
    stats->many_path_iters = h.ctl.many_iters;
    return LSSEG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched form
// ---------------------------------------------------------------------------
struct lsseg_batch {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<lsseg_ctx*> vols;
    int32_t dtype = 0;
    uint32_t cus = 256;
    uint32_t max_compact = 1024;      // resident workgroups of k_compact_batch (and above the volume count)
    uint32_t grid = 1024;             // compaction workgroups
    uint32_t sgrid = 2048;            // update and marking workgroups
    uint32_t keep_items = kKeepItemsBatch;  // traversal positions per thread of the keep step
    uint32_t compact_tile = kPosTile;       // list positions per few-path compaction tile (256 or 1024)
    bool defer_keep = true;           // LSSEG_DEFER_KEEP at creation: the keep step leaves the traversals
                                      // to the update launch
    bool deferring = false;           // the call in progress defers them (not with LSSEG_FUSE=0)
    KArgs* d_table = nullptr;
    DevCtl* d_ctls = nullptr;
    DevCtl* h_ctls = nullptr;         // pinned mirror
    std::vector<KArgs> h_table;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

lsseg_status bfail(lsseg_batch* b, lsseg_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    else g_batch_create_error = buf;
    return st;
}

#define BHIP_TRY(b, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return bfail((b), LSSEG_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

// a failure of a volume's context, reported by the batch
lsseg_status bforward(lsseg_batch* b, uint32_t v, lsseg_status st)
{
    if (st == LSSEG_OK) return st;
    return bfail(b, st, "volume %u: %s", v, b->vols[v]->err.c_str());
}

void destroy_batch_graphs(lsseg_batch* b)
{
    for (auto& g : b->graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// the batched counterparts of launch_head / launch_tail / launch_iteration: the
// update (with `fused`, in one launch with every volume's pending compaction:
// b->grid compaction workgroups, then b->sgrid update workgroups), marking and
// keep step, compaction
// The grid of a deferred batched keep step: it waits on nothing, so it runs
// wide enough that its chunks stay single sub-tiles, and the update launch
// that rebuilds the traversals gets as many work items (the compaction grid,
// about 1024 workgroups for 64 volumes, gave chunks of 6 sub-tiles and a 6%
// slower batch)
uint32_t batch_keep_grid(const lsseg_batch* b)
{
    return b->deferring ? std::max<uint32_t>(b->grid, (uint32_t)env_int("LSSEG_BATCH_KEEP_GRID", 8192)) : b->grid;
}

BatchArgs batch_args(const lsseg_batch* b)
{
    return BatchArgs{b->d_table, b->d_ctls, (uint32_t)b->vols.size(), b->keep_items, batch_keep_grid(b),
                     b->deferring ? 1u : 0u};
}

void launch_batch_head(lsseg_batch* b, int q, bool fused)
{
    const BatchArgs args = batch_args(b);
    const bool phi8 = b->vols[0]->phi8;
    if (fused) {
        uint32_t gc = b->grid;
        void* kargs[] = {(void*)&args, (void*)&q, (void*)&gc};
        (void)hipLaunchKernel(update_fused_batch_kernel(b->dtype, phi8), dim3(b->grid + b->sgrid), dim3(kBlock), kargs,
                              0, b->stream);
    } else {
        void* kargs[] = {(void*)&args, (void*)&q};
        (void)hipLaunchKernel(update_batch_kernel(b->dtype, phi8), dim3(b->sgrid), dim3(kBlock), kargs, 0, b->stream);
    }
}

void launch_batch_tail(lsseg_batch* b, int q)
{
    const BatchArgs args = batch_args(b);
    void* kargs[] = {(void*)&args, (void*)&q};
    (void)hipLaunchKernel(mark_batch_kernel(b->vols[0]->phi8), dim3(b->sgrid), dim3(kBlock), kargs, 0, b->stream);
    (void)hipLaunchKernel((const void*)&k_keep_batch, dim3(batch_keep_grid(b)), dim3(kBlock), kargs, 0, b->stream);
}

void launch_batch_compact(lsseg_batch* b, int q)
{
    const BatchArgs args = batch_args(b);
    void* kargs[] = {(void*)&args, (void*)&q};
    (void)hipLaunchKernel((const void*)&k_compact_batch, dim3(b->grid), dim3(kBlock), kargs, 0, b->stream);
}

// stamp_reserve for every volume of the batch (each has its own records and
// stamp; a volume may also have been iterated on its own)
void batch_stamp_reserve(lsseg_batch* b, uint32_t count, bool pending)
{
    for (lsseg_ctx* c : b->vols) {
        if (c->stamp == 0 || c->stamp + count - 1u > 255u) {
            hipLaunchKernelGGL(k_records_reset, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, b->stream, c->a.C,
                               c->a.g.N, c->a.ctl, pending ? 0u : 1u);
            c->stamp = 1;
        }
        c->stamp += count;
    }
}

void launch_batch_iteration(lsseg_batch* b, int q)
{
    batch_stamp_reserve(b, 1, false);
    launch_batch_head(b, q, false);
    launch_batch_tail(b, q);
    launch_batch_compact(b, q);
}

// as launch_iterations: 3 count + 1 launches, runs of kGraphIters fused
// iterations from a captured graph; LSSEG_FUSE=0 issues four per iteration
lsseg_status launch_batch_iterations(lsseg_batch* b, int count, int q)
{
    if (count <= 0) return LSSEG_OK;
    if (!env_int("LSSEG_FUSE", 1)) {
        for (int i = 0; i < count; ++i) launch_batch_iteration(b, q ^ (i & 1));
        return LSSEG_OK;
    }
    batch_stamp_reserve(b, 1, false);
    launch_batch_head(b, q, false);
    launch_batch_tail(b, q);
    q ^= 1;
    --count;
    const int chunk = env_int("LSSEG_GRAPH", 1) ? kGraphIters : 0;
    while (chunk && count >= chunk) {
        hipGraphExec_t& ge = b->graphs[q];
        if (!ge) {
            hipGraph_t g = nullptr;
            BHIP_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < chunk; ++i) {
                launch_batch_head(b, q ^ (i & 1), true);
                launch_batch_tail(b, q ^ (i & 1));
            }
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess) return bfail(b, LSSEG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ge = nullptr;
                return bfail(b, LSSEG_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
            }
        }
        batch_stamp_reserve(b, (uint32_t)chunk, true);
        BHIP_TRY(b, hipGraphLaunch(ge, b->stream));
        count -= chunk;
    }
    for (int i = 0; i < count; ++i) {
        batch_stamp_reserve(b, 1, true);
        launch_batch_head(b, q, true);
        launch_batch_tail(b, q);
        q ^= 1;
    }
    launch_batch_compact(b, q ^ 1);
    if (b->deferring) {
        const BatchArgs args = batch_args(b);
        void* kargs[] = {(void*)&args, (void*)&q};
        (void)hipLaunchKernel((const void*)&k_settle_batch, dim3(b->sgrid), dim3(kBlock), kargs, 0, b->stream);
    }
    return LSSEG_OK;
}

// Makes parity 0 the current one: swapping the two phi and list buffers (and
// their counts) is a renaming, not a copy.
lsseg_status normalise_parity(lsseg_ctx* c)
{
    if (c->q == 0) return LSSEG_OK;
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::swap(c->h_ctl->n[0], c->h_ctl->n[1]);
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl, c->h_ctl, sizeof(DevCtl), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::swap(c->a.phi[0], c->a.phi[1]);
    std::swap(c->a.list[0], c->a.list[1]);
    std::swap(c->a.S[0], c->a.S[1]);
    c->q = 0;
    destroy_graphs(c);
    return LSSEG_OK;
}

// Autotuning results, cached per (GPU architecture, source type, phi storage,
// volume count, total voxels to the next power of two, timed iterations) in
// this process and, when LSSEG_TUNING_CACHE names a file, in that file (one
// result per line, appended; read at the first autotune).
std::mutex g_tune_mu;
std::vector<std::pair<std::string, lsseg_tuning>> g_tune_cache;
bool g_tune_file_read = false;

std::string gpu_arch()
{
    hipDeviceProp_t prop{};
    int dev = 0;
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        arch = std::string(prop.gcnArchName).substr(0, std::string(prop.gcnArchName).find(':'));
    return arch;
}

std::string tuning_key(const lsseg_batch* b, int32_t iters)
{
    const std::string arch = gpu_arch();
    uint64_t voxels = 0;
    for (const lsseg_ctx* c : b->vols) voxels += c->a.g.N;
    int lg = 0;
    while ((1ull << lg) < voxels) ++lg;
    char buf[160];
    snprintf(buf, sizeof(buf), "%s/dtype%d/phi%d/volumes%zu/voxels2^%d/iters%d", arch.c_str(), b->dtype,
             b->vols[0]->phi8 ? 1 : 0, b->vols.size(), lg, iters);
    return buf;
}

bool tuning_cache_get(const std::string& key, lsseg_tuning* out)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const char* path = getenv("LSSEG_TUNING_CACHE");
    if (!g_tune_file_read && path && *path) {
        g_tune_file_read = true;
        if (FILE* f = fopen(path, "r")) {
            char k[160];
            lsseg_tuning t{};
            while (fscanf(f, "%159s %u %u %u %u %lf", k, &t.update_blocks, &t.compact_blocks, &t.keep_items,
                          &t.compact_tile, &t.ms) == 6)
                g_tune_cache.emplace_back(k, t);
            fclose(f);
        }
    }
    for (auto it = g_tune_cache.rbegin(); it != g_tune_cache.rend(); ++it)
        if (it->first == key) {
            *out = it->second;
            return true;
        }
    return false;
}

void tuning_cache_put(const std::string& key, const lsseg_tuning& t)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache.emplace_back(key, t);
    const char* path = getenv("LSSEG_TUNING_CACHE");
    if (path && *path)
        if (FILE* f = fopen(path, "a")) {
            fprintf(f, "%s %u %u %u %u %.6f\n", key.c_str(), t.update_blocks, t.compact_blocks, t.keep_items,
                    t.compact_tile, t.ms);
            fclose(f);
        }
}

lsseg_status batch_sync_and_check(lsseg_batch* b)
{
    const size_t cnt = b->vols.size();
    BHIP_TRY(b, hipMemcpyAsync(b->h_ctls, b->d_ctls, cnt * sizeof(DevCtl), hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    for (size_t v = 0; v < cnt; ++v) {
        *b->vols[v]->h_ctl = b->h_ctls[v];
        if (b->h_ctls[v].error)
            return bfail(b, LSSEG_ERR_DEVICE, "volume %zu: device error flag 0x%x (%s)", v, b->h_ctls[v].error,
                         device_error_text(b->h_ctls[v].error));
    }
    return LSSEG_OK;
}

}  // namespace

extern "C" {

const char* lsseg_batch_last_error(const lsseg_batch* b)
{
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

lsseg_status lsseg_batch_destroy(lsseg_batch* b)
{
    if (!b) return LSSEG_ERR_INVALID_ARG;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_batch_graphs(b);
    for (lsseg_ctx* c : b->vols) {
        c->batch = nullptr;
        destroy_ctx(c);
    }
    b->vols.clear();
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->d_ctls) (void)hipFree(b->d_ctls);
    if (b->h_ctls) (void)hipHostFree(b->h_ctls);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LSSEG_OK;
}

lsseg_status lsseg_batch_create(const lsseg_volume_desc* descs, const void* const* host_volumes, uint32_t count,
                                lsseg_batch** out)
{
    if (!descs || !host_volumes || !out) return bfail(nullptr, LSSEG_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (count == 0 || count > LSSEG_MAX_BATCH)
        return bfail(nullptr, LSSEG_ERR_INVALID_ARG, "batch size %u outside 1..%d", count, LSSEG_MAX_BATCH);
    for (uint32_t v = 0; v < count; ++v)
        if (descs[v].dtype != descs[0].dtype || descs[v].phi_storage != descs[0].phi_storage)
            return bfail(nullptr, LSSEG_ERR_UNSUPPORTED,
                         "volume %u has another source type or phi storage than volume 0", v);
    lsseg_batch* b = new (std::nothrow) lsseg_batch();
    if (!b) return bfail(nullptr, LSSEG_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto bail = [&](lsseg_status st) {
        g_batch_create_error = b->err;
        lsseg_batch_destroy(b);
        return st;
    };
    b->dtype = descs[0].dtype;
    b->defer_keep = env_int("LSSEG_DEFER_KEEP", 1) != 0;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
        b->err = std::string("stream/event creation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    for (uint32_t v = 0; v < count; ++v) {
        lsseg_ctx* c = nullptr;
        lsseg_status st = create_ctx(&descs[v], host_volumes[v], b->stream, &c);
        if (st != LSSEG_OK) {
            b->err = "volume " + std::to_string(v) + ": " + g_create_error;
            return bail(st);
        }
        c->batch = b;
        b->vols.push_back(c);
    }
    // one contiguous array of control blocks: the batched kernels read every
    // volume's counts at once
    if (hipMalloc((void**)&b->d_ctls, count * sizeof(DevCtl)) != hipSuccess ||
        hipMalloc((void**)&b->d_table, count * sizeof(KArgs)) != hipSuccess ||
        hipHostMalloc((void**)&b->h_ctls, count * sizeof(DevCtl), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b->err = "batch allocation failed";
        return bail(LSSEG_ERR_OUT_OF_MEMORY);
    }
    for (uint32_t v = 0; v < count; ++v) {
        lsseg_ctx* c = b->vols[v];
        if ((e = hipMemcpyAsync(&b->d_ctls[v], c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToDevice, b->stream)) !=
            hipSuccess) {
            b->err = std::string("control block copy failed: ") + hipGetErrorString(e);
            return bail(LSSEG_ERR_DEVICE);
        }
        c->a.ctl = &b->d_ctls[v];
        destroy_graphs(c);
    }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) {
        b->err = std::string("batch initialisation failed: ") + hipGetErrorString(e);
        return bail(LSSEG_ERR_DEVICE);
    }
    b->cus = b->vols[0]->cus;
    const int occ = min_occupancy({(const void*)&k_compact_batch, (const void*)&k_keep_batch});
    // the batched scans take one look-back chunk per workgroup and need more
    // workgroups than volumes
    b->max_compact = std::max<uint32_t>((uint32_t)occ * b->cus, count + 8u);
    b->grid = b->max_compact;
    b->sgrid = 8u * b->cus;
    *out = b;
    return LSSEG_OK;
}

lsseg_status lsseg_batch_count(const lsseg_batch* b, uint32_t* count)
{
    if (!b || !count) return LSSEG_ERR_INVALID_ARG;
    *count = (uint32_t)b->vols.size();
    return LSSEG_OK;
}

lsseg_status lsseg_batch_volume(lsseg_batch* b, uint32_t index, lsseg_ctx** out)
{
    if (!b || !out) return LSSEG_ERR_INVALID_ARG;
    if (index >= b->vols.size()) return bfail(b, LSSEG_ERR_INVALID_ARG, "volume index %u out of range", index);
    *out = b->vols[index];
    return LSSEG_OK;
}

lsseg_status lsseg_batch_iterate(lsseg_batch* b, const lsseg_params* params, int32_t iters, lsseg_stats* stats)
{
    if (!b || !params) return LSSEG_ERR_INVALID_ARG;
    if (iters < 0) return bfail(b, LSSEG_ERR_INVALID_ARG, "negative iteration count");
    const uint32_t cnt = (uint32_t)b->vols.size();
    lsseg_status st;
    for (uint32_t v = 0; v < cnt; ++v) {
        lsseg_ctx* c = b->vols[v];
        DevParams P;
        if ((st = bforward(b, v, make_params(c, &params[v], &P))) != LSSEG_OK) return st;
        if ((st = bforward(b, v, ensure_stats(c, (uint64_t)iters))) != LSSEG_OK) return st;
        if ((st = bforward(b, v, upload_params(c, P))) != LSSEG_OK) return st;
        if ((st = bforward(b, v, normalise_parity(c))) != LSSEG_OK) return st;
    }
    b->h_table.resize(cnt);
    // a deferred keep step leaves each volume's traversal to the next update
    // launch (the four-launch schedule places it in the keep step)
    b->deferring = b->defer_keep && env_int("LSSEG_FUSE", 1) != 0;
    for (uint32_t v = 0; v < cnt; ++v) {
        b->h_table[v] = b->vols[v]->a;
        b->h_table[v].defer = b->deferring ? 1u : 0u;
        // compaction tiles of 1024 positions unless tuned otherwise: small
        // tiles made the 64-volume batch 6% slower (more chunks over the one
        // grid all volumes share)
        const Geom& g = b->vols[v]->a.g;
        b->h_table[v].g.small_n =
            b->compact_tile == (uint32_t)kBlock ? (uint32_t)std::min<uint64_t>(g.N, 131072u) : 0u;
    }
    BHIP_TRY(b, hipMemcpyAsync(b->d_table, b->h_table.data(), cnt * sizeof(KArgs), hipMemcpyHostToDevice, b->stream));
    // reset the per-call counters (keep n[], epoch)
    BHIP_TRY(b, hipMemcpyAsync(b->h_ctls, b->d_ctls, cnt * sizeof(DevCtl), hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(b, hipStreamSynchronize(b->stream));
    for (uint32_t v = 0; v < cnt; ++v) {
        DevCtl& h = b->h_ctls[v];
        h.iter = 0;
        h.stats_cap = b->vols[v]->stats_cap;
        if (h.error) {  // as in lsseg_iterate
            h.epoch += 2;
            const KArgs& a = b->vols[v]->a;
            hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(a.g.N)), dim3(kBlock), 0, b->stream, a.keep, a.g.N,
                               (uint8_t)0);
            b->vols[v]->stamp = 0;
        }
        h.error = 0;
        h.few_iters = 0;
        h.many_iters = 0;
    }
    BHIP_TRY(b, hipMemcpyAsync(b->d_ctls, b->h_ctls, cnt * sizeof(DevCtl), hipMemcpyHostToDevice, b->stream));
    BHIP_TRY(b, hipEventRecord(b->ev0, b->stream));
    if (iters > 0 && (st = launch_batch_iterations(b, iters, 0)) != LSSEG_OK) return st;
    BHIP_TRY(b, hipGetLastError());
    BHIP_TRY(b, hipEventRecord(b->ev1, b->stream));
    if ((st = batch_sync_and_check(b)) != LSSEG_OK) return st;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    for (uint32_t v = 0; v < cnt; ++v) {
        lsseg_ctx* c = b->vols[v];
        c->q = iters & 1;
        c->last_iters = (uint64_t)iters;
    }
    if (stats) {
        for (uint32_t v = 0; v < cnt; ++v) {
            lsseg_ctx* c = b->vols[v];
            lsseg_stats& o = stats[v];
            memset(&o, 0, sizeof(o));
            o.iterations = (uint64_t)iters;
            o.device_ms = ms;
            o.active_final = c->h_ctl->n[c->q];
            o.few_path_iters = c->h_ctl->few_iters;
            o.many_path_iters = c->h_ctl->many_iters;
            if (iters > 0) {
                std::vector<uint32_t> hn(iters), hm(iters);
                std::vector<uint64_t> hc(iters);
                BHIP_TRY(b, hipMemcpy(hn.data(), c->a.n_hist, iters * sizeof(uint32_t), hipMemcpyDeviceToHost));
                BHIP_TRY(b, hipMemcpy(hc.data(), c->a.c_hist, iters * sizeof(uint64_t), hipMemcpyDeviceToHost));
                BHIP_TRY(b, hipMemcpy(hm.data(), c->a.m_hist, iters * sizeof(uint32_t), hipMemcpyDeviceToHost));
                for (int i = 0; i < iters; ++i) {
                    o.sum_active += hn[i];
                    o.sum_candidates += hc[i];
                    o.sum_new_active += hm[i];
                }
            }
        }
    }
    return LSSEG_OK;
}

lsseg_status lsseg_batch_checkpoint_save(lsseg_batch* b, int32_t slot)
{
    if (!b) return LSSEG_ERR_INVALID_ARG;
    for (uint32_t v = 0; v < b->vols.size(); ++v) {
        lsseg_status st = bforward(b, v, lsseg_checkpoint_save(b->vols[v], slot));
        if (st != LSSEG_OK) return st;
    }
    return LSSEG_OK;
}

lsseg_status lsseg_batch_checkpoint_restore(lsseg_batch* b, int32_t slot)
{
    if (!b) return LSSEG_ERR_INVALID_ARG;
    for (uint32_t v = 0; v < b->vols.size(); ++v) {
        lsseg_status st = bforward(b, v, lsseg_checkpoint_restore(b->vols[v], slot));
        if (st != LSSEG_OK) return st;
    }
    return LSSEG_OK;
}

lsseg_status lsseg_batch_get_tuning(const lsseg_batch* b, lsseg_tuning* out)
{
    if (!b || !out) return LSSEG_ERR_INVALID_ARG;
    memset(out, 0, sizeof(*out));
    out->update_blocks = b->sgrid;
    out->compact_blocks = b->grid;
    out->keep_items = b->keep_items;
    out->compact_tile = b->compact_tile;
    return LSSEG_OK;
}

lsseg_status lsseg_batch_set_tuning(lsseg_batch* b, const lsseg_tuning* t)
{
    if (!b || !t) return LSSEG_ERR_INVALID_ARG;
    const uint32_t min_compact = (uint32_t)b->vols.size() + 1u;
    if (t->update_blocks == 0 || t->compact_blocks < min_compact || t->compact_blocks > b->max_compact)
        return bfail(b, LSSEG_ERR_INVALID_ARG, "launch shape %u/%u outside 1.. / %u..%u", t->update_blocks,
                     t->compact_blocks, min_compact, b->max_compact);
    const uint32_t ki = t->keep_items ? t->keep_items : b->keep_items;
    const uint32_t ct = t->compact_tile ? t->compact_tile : b->compact_tile;
    if ((ki != 1 && ki != 2) || (ct != (uint32_t)kBlock && ct != (uint32_t)kPosTile))
        return bfail(b, LSSEG_ERR_INVALID_ARG, "keep items %u (1 or 2) or compaction tile %u (%d or %d)", ki, ct,
                     kBlock, kPosTile);
    if (t->update_blocks != b->sgrid || t->compact_blocks != b->grid || ki != b->keep_items || ct != b->compact_tile)
        destroy_batch_graphs(b);
    b->sgrid = t->update_blocks;
    b->grid = t->compact_blocks;
    b->keep_items = ki;
    b->compact_tile = ct;
    return LSSEG_OK;
}

lsseg_status lsseg_batch_autotune(lsseg_batch* b, const lsseg_params* params, int32_t iters, lsseg_tuning* out)
{
    if (!b || !params) return LSSEG_ERR_INVALID_ARG;
    if (iters <= 0) return bfail(b, LSSEG_ERR_INVALID_ARG, "autotune needs a positive iteration count");
    const std::string key = tuning_key(b, iters);
    lsseg_tuning best{};
    if (!env_int("LSSEG_AUTOTUNE_FORCE", 0) && tuning_cache_get(key, &best)) {
        best.cached = 1;
        best.candidates = 0;
        lsseg_status st = lsseg_batch_set_tuning(b, &best);
        if (st == LSSEG_OK && out) *out = best;
        return st;
    }
    const int32_t slot = LSSEG_MAX_CHECKPOINTS - 1;
    lsseg_status st = lsseg_batch_checkpoint_save(b, slot);
    if (st != LSSEG_OK) return st;
    uint32_t timed = 0;
    // device time of `iters` iterations from the saved state with tuning t;
    // the first run captures the graphs, the second is timed
    auto time_shape = [&](const lsseg_tuning& t, double* ms) -> lsseg_status {
        lsseg_status s2 = lsseg_batch_set_tuning(b, &t);
        if (s2 != LSSEG_OK) return s2;
        for (int rep = 0; rep < 2; ++rep) {
            if ((s2 = lsseg_batch_checkpoint_restore(b, slot)) != LSSEG_OK) return s2;
            std::vector<lsseg_stats> sts(b->vols.size());
            if ((s2 = lsseg_batch_iterate(b, params, iters, sts.data())) != LSSEG_OK) return s2;
            *ms = sts[0].device_ms;
        }
        ++timed;
        return LSSEG_OK;
    };
    // coordinate descent: the launch shape (workgroups of the update and
    // marking launches, of the keep and compaction launches) at the default
    // widths, then the keep step's positions per thread, then the compaction's
    // tile of list positions
    const uint32_t occ = b->max_compact / b->cus;
    std::vector<lsseg_tuning> cands;
    for (uint32_t su : {2u, 4u, 8u, 16u})
        for (uint32_t cd : {occ, (occ + 1) / 2, 1u}) {
            lsseg_tuning t{};
            t.update_blocks = su * b->cus;
            t.compact_blocks = std::max<uint32_t>(std::max(1u, cd) * b->cus, (uint32_t)b->vols.size() + 8u);
            t.keep_items = kKeepItemsBatch;
            t.compact_tile = kPosTile;
            bool dup = false;
            for (auto& o : cands) dup |= o.update_blocks == t.update_blocks && o.compact_blocks == t.compact_blocks;
            if (!dup) cands.push_back(t);
        }
    lsseg_tuning start{};
    (void)lsseg_batch_get_tuning(b, &start);
    // a failed timing run leaves the caller's tuning and state as they were
    auto give_up = [&](lsseg_status s2) {
        (void)lsseg_batch_set_tuning(b, &start);
        (void)lsseg_batch_checkpoint_restore(b, slot);
        return s2;
    };
    best.ms = 1e30;
    for (auto& t : cands) {
        double ms = 0;
        if ((st = time_shape(t, &ms)) != LSSEG_OK) return give_up(st);
        if (ms < best.ms) {
            best = t;
            best.ms = ms;
        }
    }
    for (int stage = 0; stage < 2; ++stage) {
        lsseg_tuning t = best;
        if (stage == 0) t.keep_items = best.keep_items == 1 ? 2 : 1;
        else t.compact_tile = best.compact_tile == (uint32_t)kBlock ? (uint32_t)kPosTile : (uint32_t)kBlock;
        double ms = 0;
        if ((st = time_shape(t, &ms)) != LSSEG_OK) return give_up(st);
        if (ms < best.ms) {
            best = t;
            best.ms = ms;
        }
    }
    best.candidates = timed;
    best.cached = 0;
    if ((st = lsseg_batch_set_tuning(b, &best)) != LSSEG_OK) return st;
    if ((st = lsseg_batch_checkpoint_restore(b, slot)) != LSSEG_OK) return st;
    tuning_cache_put(key, best);
    if (out) *out = best;
    return LSSEG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// launch shapes and widths of a single context
// ---------------------------------------------------------------------------
namespace {

std::string ctx_tuning_key(const lsseg_ctx* c, int32_t iters)
{
    int lg = 0;
    while ((1ull << lg) < c->a.g.N) ++lg;
    char buf[160];
    snprintf(buf, sizeof(buf), "%s/single/dtype%d/phi%d/voxels2^%d/iters%d", gpu_arch().c_str(), c->dtype,
             c->phi8 ? 1 : 0, lg, iters);
    return buf;
}

}  // namespace

extern "C" {

lsseg_status lsseg_get_tuning(const lsseg_ctx* c, lsseg_tuning* out)
{
    if (!c || !out) return LSSEG_ERR_INVALID_ARG;
    *out = lsseg_tuning{};
    out->update_blocks = c->sgrid;
    out->compact_blocks = c->grid;
    out->keep_items = c->keep_fn == (const void*)&k_keep<2> ? 2u : 1u;
    out->compact_tile = c->a.g.small_n > 0 ? (uint32_t)kBlock : (uint32_t)kPosTile;
    return LSSEG_OK;
}

lsseg_status lsseg_set_tuning(lsseg_ctx* c, const lsseg_tuning* t)
{
    if (!c || !t) return LSSEG_ERR_INVALID_ARG;
    lsseg_tuning cur{};
    (void)lsseg_get_tuning(c, &cur);
    const uint32_t ub = t->update_blocks ? t->update_blocks : cur.update_blocks;
    const uint32_t cb = t->compact_blocks ? t->compact_blocks : cur.compact_blocks;
    const uint32_t ki = t->keep_items ? t->keep_items : cur.keep_items;
    const uint32_t ct = t->compact_tile ? t->compact_tile : cur.compact_tile;
    if (ub < 8 || (ub & 7u) || ub > (1u << 20) || cb < 16 || (cb & 7u) || cb > (1u << 20))
        return fail(c, LSSEG_ERR_INVALID_ARG,
                    "launch shape %u/%u: workgroup counts are multiples of 8, at least 8 (update) and 16 (scan), "
                    "at most 2^20", ub, cb);
    if ((ki != 1 && ki != 2) || (ct != (uint32_t)kBlock && ct != (uint32_t)kPosTile))
        return fail(c, LSSEG_ERR_INVALID_ARG, "keep items %u (1 or 2) or compaction tile %u (%d or %d)", ki, ct,
                    kBlock, kPosTile);
    if (ct == (uint32_t)kBlock && c->small_n_max == 0)
        return fail(c, LSSEG_ERR_INVALID_ARG,
                    "compaction tiles of %d positions need a context created with them (LSSEG_SMALL_TILE_N > 0)",
                    kBlock);
    if (ub == cur.update_blocks && cb == cur.compact_blocks && ki == cur.keep_items && ct == cur.compact_tile)
        return LSSEG_OK;
    destroy_graphs(c);  // captured launch shapes and kernel arguments change
    c->sgrid = ub;
    c->grid = cb;
    c->a.fused_compact_max = (cb / 4u * 3u) & ~7u;
    c->fgrid = cb;
    c->keep_fn = ki == 2 ? (const void*)&k_keep<2> : (const void*)&k_keep<1>;
    c->a.g.small_n = ct == (uint32_t)kBlock ? c->small_n_max : 0u;
    return LSSEG_OK;
}

lsseg_status lsseg_autotune(lsseg_ctx* c, const lsseg_params* params, int32_t iters, lsseg_tuning* out)
{
    if (!c || !params) return LSSEG_ERR_INVALID_ARG;
    if (iters <= 0) return fail(c, LSSEG_ERR_INVALID_ARG, "autotune needs a positive iteration count");
    const std::string key = ctx_tuning_key(c, iters);
    lsseg_tuning best{};
    if (!env_int("LSSEG_AUTOTUNE_FORCE", 0) && tuning_cache_get(key, &best)) {
        if (best.compact_tile == (uint32_t)kBlock && c->small_n_max == 0) best.compact_tile = kPosTile;
        best.cached = 1;
        best.candidates = 0;
        lsseg_status st = lsseg_set_tuning(c, &best);
        if (st == LSSEG_OK && out) *out = best;
        return st;
    }
    const int32_t slot = LSSEG_MAX_CHECKPOINTS - 1;
    lsseg_status st = lsseg_checkpoint_save(c, slot);
    if (st != LSSEG_OK) return st;
    uint32_t timed = 0;
    // device time of `iters` iterations from the saved state with tuning t;
    // the first run captures the graphs, the second is timed
    auto time_shape = [&](const lsseg_tuning& t, double* ms) -> lsseg_status {
        lsseg_status s2 = lsseg_set_tuning(c, &t);
        if (s2 != LSSEG_OK) return s2;
        for (int rep = 0; rep < 2; ++rep) {
            if ((s2 = lsseg_checkpoint_restore(c, slot)) != LSSEG_OK) return s2;
            lsseg_stats stt{};
            if ((s2 = lsseg_iterate(c, params, iters, &stt)) != LSSEG_OK) return s2;
            *ms = stt.device_ms;
        }
        ++timed;
        return LSSEG_OK;
    };
    // coordinate descent: the launch shape (update and marking workgroups,
    // keep and compaction workgroups) at the current widths, then the keep
    // step's positions per thread, then the compaction's tile
    lsseg_tuning start{};
    (void)lsseg_get_tuning(c, &start);
    std::vector<lsseg_tuning> cands;
    const uint32_t half = std::max<uint32_t>(16u, (c->grid0 / 2u) & ~7u);
    const uint32_t two = std::max<uint32_t>(16u, (2u * c->cus) & ~7u);
    // scan grids above the resident count too: the keep step of a deferred
    // traversal waits on nothing, and the vessels 512^3 run measured 2% faster
    // at 1.5 and 2 times the occupancy grid
    const uint32_t more = (c->grid0 * 3u / 2u) & ~7u, twice = (c->grid0 * 2u) & ~7u;
    for (uint32_t su : {4u, 8u, 16u})
        for (uint32_t cb : {c->grid0, half, two, more, twice}) {
            lsseg_tuning t = start;
            t.update_blocks = std::max<uint32_t>(8u, (su * c->cus) & ~7u);
            t.compact_blocks = cb;
            bool dup = false;
            for (auto& o : cands) dup |= o.update_blocks == t.update_blocks && o.compact_blocks == t.compact_blocks;
            if (!dup) cands.push_back(t);
        }
    // a failed timing run leaves the caller's tuning and state as they were
    auto give_up = [&](lsseg_status s2) {
        (void)lsseg_set_tuning(c, &start);
        (void)lsseg_checkpoint_restore(c, slot);
        return s2;
    };
    best.ms = 1e30;
    for (auto& t : cands) {
        double ms = 0;
        if ((st = time_shape(t, &ms)) != LSSEG_OK) return give_up(st);
        if (ms < best.ms) {
            best = t;
            best.ms = ms;
        }
    }
    for (int stage = 0; stage < 2; ++stage) {
        lsseg_tuning t = best;
        if (stage == 0) t.keep_items = best.keep_items == 1 ? 2 : 1;
        else if (c->small_n_max > 0)
            t.compact_tile = best.compact_tile == (uint32_t)kBlock ? (uint32_t)kPosTile : (uint32_t)kBlock;
        else
            continue;
        double ms = 0;
        if ((st = time_shape(t, &ms)) != LSSEG_OK) return give_up(st);
        if (ms < best.ms) {
            best = t;
            best.ms = ms;
        }
    }
    best.candidates = timed;
    best.cached = 0;
    if ((st = lsseg_set_tuning(c, &best)) != LSSEG_OK) return st;
    if ((st = lsseg_checkpoint_restore(c, slot)) != LSSEG_OK) return st;
    tuning_cache_put(key, best);
    if (out) *out = best;
    return LSSEG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// differentiable form
// ---------------------------------------------------------------------------
struct lsseg_record {
    lsseg_ctx* ctx = nullptr;
    int32_t iters = 0;
    lsseg_params params{};
    DevParams* d_prm = nullptr;       // the recorded run's parameters
    DevParams* d_prm_replay = nullptr;
    std::vector<uint32_t> n;          // n_k
    std::vector<uint64_t> off;        // offsets of V_k in lists / dec
    uint64_t total = 0;
    uint64_t* offs = nullptr;         // the same offsets, written by the recording run
    uint64_t cap = 0;                 // entries that lists / dec hold
    void* lists_block = nullptr;      // lists, undo and dec
    uint32_t* lists = nullptr;
    float* undo = nullptr;            // per entry the write buffer's value its update overwrote
    uint8_t* dec = nullptr;
    float* fin[2] = {nullptr, nullptr};  // (read, write) buffers after the last iteration
    float* work[2] = {nullptr, nullptr};
    float* adj[2] = {nullptr, nullptr};
    float* lin = nullptr;             // unpadded staging
    float* d_theta = nullptr;
    std::vector<void*> allocs;
};

namespace {

constexpr int32_t kRecordSlot = LSSEG_MAX_CHECKPOINTS - 1;

// A destroyed context leaves its records detached: they keep their memory
// until lsseg_record_destroy and refuse every other call.
void detach_records(lsseg_ctx* c)
{
    for (lsseg_record* r : c->records) r->ctx = nullptr;
    c->records.clear();
}

lsseg_status ralloc(lsseg_ctx* c, lsseg_record* r, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, LSSEG_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    r->allocs.push_back(*p);
    return LSSEG_OK;
}

#define LSSEG_DTYPE_KERNEL(name)                                                          \
    const void* name##_kernel(int32_t dtype)                                              \
    {                                                                                     \
        switch (dtype) {                                                                  \
        case LSSEG_U8: return (const void*)&name<uint8_t, false>;                         \
        case LSSEG_I8: return (const void*)&name<int8_t, false>;                          \
        case LSSEG_U16: return (const void*)&name<uint16_t, false>;                       \
        case LSSEG_I16: return (const void*)&name<int16_t, false>;                        \
        case LSSEG_U32: return (const void*)&name<uint32_t, false>;                       \
        case LSSEG_I32: return (const void*)&name<int32_t, false>;                        \
        default: return (const void*)&name<float, true>;                                  \
        }                                                                                 \
    }
LSSEG_DTYPE_KERNEL(k_replay)
LSSEG_DTYPE_KERNEL(k_adjoint)
#undef LSSEG_DTYPE_KERNEL

uint32_t list_grid(uint64_t n) { return dense_grid(std::max<uint64_t>(n, 1)); }

void launch_replay(lsseg_ctx* c, lsseg_record* r, const KArgs& ra, int k, const float* R, float* W)
{
    const uint32_t* V = r->lists + r->off[k];
    const uint8_t* D = r->dec + r->off[k];
    uint32_t n = r->n[k];
    void* args[] = {(void*)&ra, (void*)&V, (void*)&D, (void*)&n, (void*)&R, (void*)&W};
    (void)hipLaunchKernel(k_replay_kernel(c->dtype), dim3(list_grid(n)), dim3(kBlock), args, 0, c->stream);
}

// the context's geometry and source with the record's parameters
KArgs record_args(lsseg_ctx* c, const DevParams* prm)
{
    KArgs ra = c->a;
    ra.prm = prm;
    return ra;
}

lsseg_status check_record(lsseg_ctx* c, const lsseg_record* r)
{
    if (!c || !r) return LSSEG_ERR_INVALID_ARG;
    if (r->ctx != c) return fail(c, LSSEG_ERR_STATE, "the record belongs to another (or a destroyed) context");
    return LSSEG_OK;
}

// offsets of the record's buffers in its one allocation, 256-byte aligned
struct RecordLayout {
    size_t end = 0;
    size_t take(size_t bytes)
    {
        const size_t o = end;
        end += (bytes + 255) / 256 * 256;
        return o;
    }
};

// bound of the first guess of sum_k n_k without an earlier record (5 bytes each)
constexpr uint64_t kRecordFirstGuess = 1ull << 27;

// lists, undo values and decisions of `cap` entries (9 bytes each), replacing
// any earlier block
lsseg_status record_lists_alloc(lsseg_ctx* c, lsseg_record* r, uint64_t cap)
{
    if (r->lists_block) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        (void)hipFree(r->lists_block);
        r->allocs.erase(std::remove(r->allocs.begin(), r->allocs.end(), r->lists_block), r->allocs.end());
        r->lists_block = nullptr;
    }
    const size_t lb = (size_t)(cap * sizeof(uint32_t) + 255) / 256 * 256;
    void* p = nullptr;
    lsseg_status st = ralloc(c, r, &p, 2 * lb + cap);
    if (st != LSSEG_OK) return st;
    r->lists_block = p;
    r->lists = (uint32_t*)p;
    r->undo = (float*)((uint8_t*)p + lb);
    r->dec = (uint8_t*)p + 2 * lb;
    r->cap = cap;
    return LSSEG_OK;
}

// the update launch of a recording run (FP32 phi): the decisions of every
// update go to bits 1-7 of its temporal-flag byte, the value it overwrites to
// the record's undo slots
const void* update_record_kernel(int32_t dtype, bool fused)
{
    switch (dtype) {
    case LSSEG_U8: return fused ? (const void*)&k_update_fused_record<uint8_t, false>
                                : (const void*)&k_update_record<uint8_t, false>;
    case LSSEG_I8: return fused ? (const void*)&k_update_fused_record<int8_t, false>
                                : (const void*)&k_update_record<int8_t, false>;
    case LSSEG_U16: return fused ? (const void*)&k_update_fused_record<uint16_t, false>
                                 : (const void*)&k_update_record<uint16_t, false>;
    case LSSEG_I16: return fused ? (const void*)&k_update_fused_record<int16_t, false>
                                 : (const void*)&k_update_record<int16_t, false>;
    case LSSEG_U32: return fused ? (const void*)&k_update_fused_record<uint32_t, false>
                                 : (const void*)&k_update_record<uint32_t, false>;
    case LSSEG_I32: return fused ? (const void*)&k_update_fused_record<int32_t, false>
                                 : (const void*)&k_update_record<int32_t, false>;
    default: return fused ? (const void*)&k_update_fused_record<float, true>
                          : (const void*)&k_update_record<float, true>;
    }
}

// the fused update launch of a recording run with the deferred keep step
const void* update_record_defer_kernel(int32_t dtype)
{
    switch (dtype) {
    case LSSEG_U8: return (const void*)&k_update_fused_defer_record<uint8_t, false>;
    case LSSEG_I8: return (const void*)&k_update_fused_defer_record<int8_t, false>;
    case LSSEG_U16: return (const void*)&k_update_fused_defer_record<uint16_t, false>;
    case LSSEG_I16: return (const void*)&k_update_fused_defer_record<int16_t, false>;
    case LSSEG_U32: return (const void*)&k_update_fused_defer_record<uint32_t, false>;
    case LSSEG_I32: return (const void*)&k_update_fused_defer_record<int32_t, false>;
    default: return (const void*)&k_update_fused_defer_record<float, true>;
    }
}

// One recording run of `iters` iterations from the context's state: the
// launches of lsseg_iterate (each compaction in the next update's launch).
// The update keeps each voxel's decisions beside its temporal flag and the
// value it overwrites in the record, and the marking writes V_k in traversal
// order with its decisions at the device offsets r->offs.  The final buffer
// pair is kept: the backward pass steps back from it.
lsseg_status record_run(lsseg_ctx* c, lsseg_record* r, int32_t iters)
{
    const uint64_t N = c->a.g.N;
    HIP_TRY(c, hipMemsetAsync(r->offs, 0, sizeof(uint64_t), c->stream));
    // reset the per-call counters (keep n[], epoch)
    HIP_TRY(c, hipMemcpyAsync(c->h_ctl, c->a.ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->h_ctl->iter = 0;
    c->h_ctl->stats_cap = c->stats_cap;
    // after a call that ended with a device error, look-back states of its
    // unfinished scans may carry the current epoch (skip past it), and an
    // unfinished compaction may have left keep bytes set (clear them)
    if (c->h_ctl->error) {
        c->h_ctl->epoch += 2;
        hipLaunchKernelGGL(k_fill_u8, dim3(dense_grid(c->a.g.N)), dim3(kBlock), 0, c->stream, c->a.keep, c->a.g.N,
                           (uint8_t)0);
        c->stamp = 0;  // the records are cleared before the next marking
    }
    c->h_ctl->error = 0;
    c->h_ctl->few_iters = 0;
    c->h_ctl->many_iters = 0;
    HIP_TRY(c, hipMemcpyAsync(c->a.ctl, c->h_ctl, sizeof(DevCtl), hipMemcpyHostToDevice, c->stream));
    int q = c->q;
    const RecordOut ro{r->offs, r->cap, r->lists, r->dec, r->undo};
    // with the deferred keep step each fused update launch rebuilds the
    // traversal and records by the rebuilt positions, as lsseg_iterate runs it
    const bool defer = c->defer_keep;
    c->a.defer = defer ? 1u : 0u;
    c->a.keep_grid = defer_keep_grid(c);
    c->a.keep_items = c->keep_fn == (const void*)&k_keep<2> ? 2u : 1u;
    const void* keep_fn = !defer ? c->keep_fn
                          : c->a.keep_items == 2 ? (const void*)&k_keep_defer<2> : (const void*)&k_keep_defer<1>;
    for (int k = 0; k < iters; ++k) {
        stamp_reserve(c, 1, k > 0);
        void* args[] = {(void*)&c->a, (void*)&q, (void*)&ro};
        const void* head = k == 0 ? update_record_kernel(c->dtype, false)
                           : defer ? update_record_defer_kernel(c->dtype) : update_record_kernel(c->dtype, true);
        (void)hipLaunchKernel(head, dim3(k > 0 ? c->grid : c->sgrid), dim3(kBlock), args, 0, c->stream);
        (void)hipLaunchKernel((const void*)&k_mark_record, dim3(c->sgrid), dim3(kBlock), args, 0, c->stream);
        launch(c, keep_fn, defer ? c->a.keep_grid : c->grid, q);
        q ^= 1;
    }
    if (iters > 0) {
        launch(c, (const void*)&k_compact<kCompact>, c->grid, q ^ 1);
        if (defer) hipLaunchKernelGGL(k_settle_traversal, dim3(c->sgrid), dim3(kBlock), 0, c->stream, c->a, q);
    }
    c->a.defer = 0;
    HIP_TRY(c, hipMemcpyAsync(r->fin[0], c->a.phi[q], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(r->fin[1], c->a.phi[q ^ 1], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipGetLastError());
    lsseg_status st = sync_and_check(c);
    if (st != LSSEG_OK) return st;
    c->q = q;
    c->last_iters = (uint64_t)iters;
    return LSSEG_OK;
}

}  // namespace

extern "C" {

lsseg_status lsseg_record_destroy(lsseg_record* r)
{
    if (!r) return LSSEG_ERR_INVALID_ARG;
    if (r->ctx) {
        if (r->ctx->stream) (void)hipStreamSynchronize(r->ctx->stream);
        auto& v = r->ctx->records;
        v.erase(std::remove(v.begin(), v.end(), r), v.end());
    }
    for (void* p : r->allocs) (void)hipFree(p);
    delete r;
    return LSSEG_OK;
}

lsseg_status lsseg_fwd_record(lsseg_ctx* c, const lsseg_params* prm, int32_t iters, lsseg_record** out)
{
    if (!c || !prm || !out) return LSSEG_ERR_INVALID_ARG;
    *out = nullptr;
    if (iters <= 0) return fail(c, LSSEG_ERR_INVALID_ARG, "the iteration count must be positive");
    if (c->phi8) return fail(c, LSSEG_ERR_UNSUPPORTED, "the differentiable form needs FP32 phi storage");
    DevParams P;
    lsseg_status st = make_params(c, prm, &P);
    if (st != LSSEG_OK) return st;
    const uint64_t N = c->a.g.N;
    if ((st = ensure_stats(c, (uint64_t)iters)) != LSSEG_OK) return st;
    if ((st = upload_params(c, P)) != LSSEG_OK) return st;
    // the start state, for a second run if the first guess of the size was short
    if ((st = lsseg_checkpoint_save(c, kRecordSlot)) != LSSEG_OK) return st;
    lsseg_record* r = new (std::nothrow) lsseg_record();
    if (!r) return fail(c, LSSEG_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto bail = [&](lsseg_status e) {
        lsseg_record_destroy(r);
        return e;
    };
    r->ctx = c;
    r->iters = iters;
    r->params = *prm;
    // everything but the lists in one allocation
    const size_t NF = N * sizeof(float);
    RecordLayout lay;
    const size_t o_prm = lay.take(sizeof(DevParams)), o_prm_replay = lay.take(sizeof(DevParams));
    const size_t o_theta = lay.take(kCountSlots * kThetaStride * sizeof(float));
    const size_t o_offs = lay.take(((size_t)iters + 1) * sizeof(uint64_t));
    const size_t o_fin = lay.take(2 * NF);
    const size_t o_work = lay.take(4 * NF), o_lin = lay.take((size_t)c->nx * c->ny * c->nz * sizeof(float));
    void* p = nullptr;
    if ((st = ralloc(c, r, &p, lay.end)) != LSSEG_OK) return bail(st);
    char* base = static_cast<char*>(p);
    r->d_prm = (DevParams*)(base + o_prm);
    r->d_prm_replay = (DevParams*)(base + o_prm_replay);
    r->d_theta = (float*)(base + o_theta);  // per-workgroup slots of dL/dT, dL/deps, dL/dalpha
    r->offs = (uint64_t*)(base + o_offs);
    for (int i = 0; i < 2; ++i) {
        r->fin[i] = (float*)(base + o_fin + (size_t)i * NF);
        r->work[i] = (float*)(base + o_work + (size_t)i * NF);
        r->adj[i] = (float*)(base + o_work + (size_t)(2 + i) * NF);
    }
    r->lin = (float*)(base + o_lin);
    HIP_TRY(c, hipMemcpyAsync(r->d_prm, &P, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
    // first guess of sum_k n_k: the last record of this context, scaled to
    // `iters`, else an eighth of the volume per iteration (bounded)
    uint64_t cap = c->record_iters > 0
                       ? c->record_total * (uint64_t)iters / (uint64_t)c->record_iters * 5 / 4 + 4096
                       : std::min<uint64_t>((uint64_t)iters * (N / 8 + 1), kRecordFirstGuess);
    const int force = env_int("LSSEG_RECORD_CAPACITY", 0);
    if (force > 0) cap = (uint64_t)force;
    if ((st = record_lists_alloc(c, r, cap)) != LSSEG_OK) return bail(st);
    if ((st = record_run(c, r, iters)) != LSSEG_OK) return bail(st);
    r->off.resize((size_t)iters + 1);
    HIP_TRY(c, hipMemcpy(r->off.data(), r->offs, r->off.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (r->off[iters] > r->cap) {
        // short: record again from the start with the exact size
        if ((st = record_lists_alloc(c, r, r->off[iters])) != LSSEG_OK) return bail(st);
        if ((st = lsseg_checkpoint_restore(c, kRecordSlot)) != LSSEG_OK) return bail(st);
        if ((st = record_run(c, r, iters)) != LSSEG_OK) return bail(st);
        HIP_TRY(c, hipMemcpy(r->off.data(), r->offs, r->off.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (r->off[iters] > r->cap)
            return bail(fail(c, LSSEG_ERR_DEVICE, "the second recording run differs from the first"));
    }
    r->n.resize(iters);
    for (int k = 0; k < iters; ++k) r->n[k] = (uint32_t)(r->off[k + 1] - r->off[k]);
    r->total = r->off[iters];
    c->record_total = r->total;
    c->record_iters = iters;
    c->records.push_back(r);
    *out = r;
    return LSSEG_OK;
}

lsseg_status lsseg_record_info(const lsseg_record* r, int32_t* iters, uint64_t* total_active)
{
    if (!r) return LSSEG_ERR_INVALID_ARG;
    if (iters) *iters = r->iters;
    if (total_active) *total_active = r->total;
    return LSSEG_OK;
}

lsseg_status lsseg_record_get(const lsseg_record* r, uint32_t* n_per_iter, uint32_t* lists, uint8_t* decisions)
{
    if (!r || !r->ctx) return LSSEG_ERR_INVALID_ARG;
    lsseg_ctx* c = r->ctx;
    if (n_per_iter) memcpy(n_per_iter, r->n.data(), r->n.size() * sizeof(uint32_t));
    if (lists && r->total) {
        std::vector<uint32_t> h(r->total);
        HIP_TRY(c, hipMemcpy(h.data(), r->lists, r->total * sizeof(uint32_t), hipMemcpyDeviceToHost));
        taus_of_ids(h.data(), r->total);
        HIP_TRY(c, hipMemcpy(lists, h.data(), r->total * sizeof(uint32_t), hipMemcpyDefault));
    }
    if (decisions && r->total) HIP_TRY(c, hipMemcpy(decisions, r->dec, r->total, hipMemcpyDefault));
    return LSSEG_OK;
}

lsseg_status lsseg_fwd_replay(lsseg_ctx* c, lsseg_record* r, const float* phi0, const lsseg_params* prm,
                              float* phi_out)
{
    lsseg_status st = check_record(c, r);
    if (st != LSSEG_OK) return st;
    if (!phi0 || !prm || !phi_out) return fail(c, LSSEG_ERR_INVALID_ARG, "null argument");
    DevParams P;
    if ((st = make_params(c, prm, &P)) != LSSEG_OK) return st;
    const uint64_t N = c->a.g.N, cnt = (uint64_t)c->nx * c->ny * c->nz;
    HIP_TRY(c, hipMemcpyAsync(r->d_prm_replay, &P, sizeof(DevParams), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(r->lin, phi0, cnt * sizeof(float), hipMemcpyDefault, c->stream));
    hipLaunchKernelGGL(k_lin_to_tiled, dim3(dense_grid(N)), dim3(kBlock), 0, c->stream, c->a.g, r->lin, r->work[0],
                       c->nx, c->ny, c->nz, 1.0f);
    HIP_TRY(c, hipMemcpyAsync(r->work[1], r->work[0], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    const KArgs ra = record_args(c, r->d_prm_replay);
    int rd = 0;
    for (int k = 0; k < r->iters; ++k) {
        launch_replay(c, r, ra, k, r->work[rd], r->work[rd ^ 1]);
        rd ^= 1;
    }
    const float* none = nullptr;
    hipLaunchKernelGGL(k_tiled_sum_to_lin, dim3(dense_grid(cnt)), dim3(kBlock), 0, c->stream, c->a.g, r->work[rd],
                       none, r->lin, c->nx, c->ny, c->nz);
    HIP_TRY(c, hipMemcpyAsync(phi_out, r->lin, cnt * sizeof(float), hipMemcpyDefault, c->stream));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LSSEG_OK;
}

lsseg_status lsseg_bwd(lsseg_ctx* c, lsseg_record* r, const float* dphi, float* dphi0, double* dtheta)
{
    lsseg_status st = check_record(c, r);
    if (st != LSSEG_OK) return st;
    if (!dphi || !dphi0) return fail(c, LSSEG_ERR_INVALID_ARG, "null argument");
    const uint64_t N = c->a.g.N, cnt = (uint64_t)c->nx * c->ny * c->nz;
    const uint32_t dg = dense_grid(N);
    // Abar_K = dL/dphi (padding 0), Bbar_K = 0
    HIP_TRY(c, hipMemcpyAsync(r->lin, dphi, cnt * sizeof(float), hipMemcpyDefault, c->stream));
    hipLaunchKernelGGL(k_lin_to_tiled, dim3(dg), dim3(kBlock), 0, c->stream, c->a.g, r->lin, r->adj[0], c->nx, c->ny,
                       c->nz, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(dg), dim3(kBlock), 0, c->stream, r->adj[1], N, 0.0f);
    HIP_TRY(c, hipMemsetAsync(r->d_theta, 0, kCountSlots * kThetaStride * sizeof(float), c->stream));
    const KArgs ra = record_args(c, r->d_prm);
    int ab = 0;  // adj[ab] = Abar, adj[ab ^ 1] = Bbar
    // From the final pair (A_K, B_K) backwards: iteration k read A_k = B_{k+1}
    // and turned B_k into A_{k+1} at V_k, so the adjoint launch of iteration k
    // reads A_k and puts the recorded values of V_k back into the other
    // buffer, which then holds B_k = A_{k-1}.  Iteration 0's write buffer is
    // never read again.
    HIP_TRY(c, hipMemcpyAsync(r->work[0], r->fin[0], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(r->work[1], r->fin[1], N * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    int rd = 1;  // A_{K-1} = B_K
    for (int k = r->iters - 1; k >= 0; --k) {
        const uint32_t* V = r->lists + r->off[k];
        const uint8_t* D = r->dec + r->off[k];
        uint32_t n = r->n[k];
        const float* A = r->work[rd];
        float* X = r->adj[ab];
        float* Y = r->adj[ab ^ 1];
        float* th = r->d_theta;
        float* Bres = r->work[rd ^ 1];
        const float* U = k > 0 ? r->undo + r->off[k] : nullptr;
        void* args[] = {(void*)&ra, (void*)&V, (void*)&D, (void*)&n, (void*)&A,
                        (void*)&X, (void*)&Y, (void*)&th, (void*)&Bres, (void*)&U};
        (void)hipLaunchKernel(k_adjoint_kernel(c->dtype), dim3(list_grid(n)), dim3(kBlock), args, 0, c->stream);
        ab ^= 1;
        rd ^= 1;
    }
    // phi0 fed both buffers: its adjoint is Abar_0 + Bbar_0
    hipLaunchKernelGGL(k_tiled_sum_to_lin, dim3(dense_grid(cnt)), dim3(kBlock), 0, c->stream, c->a.g, r->adj[0],
                       r->adj[1], r->lin, c->nx, c->ny, c->nz);
    HIP_TRY(c, hipMemcpyAsync(dphi0, r->lin, cnt * sizeof(float), hipMemcpyDefault, c->stream));
    std::vector<float> slots((size_t)kCountSlots * kThetaStride);
    HIP_TRY(c, hipMemcpyAsync(slots.data(), r->d_theta, slots.size() * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // the slots' FP32 sums, added pairwise (returned widened to double)
    for (int j = 0; j < 3; ++j) {
        float col[kCountSlots];
        for (int i = 0; i < kCountSlots; ++i) col[i] = slots[(size_t)i * kThetaStride + j];
        if (dtheta) dtheta[j] = (double)pairwise_sum(col, kCountSlots);
    }
    return LSSEG_OK;
}

}  // extern "C"
