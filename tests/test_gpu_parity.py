"""Parity of the HIP engine with the oracle, through the C ABI (needs a GPU).

Bar: bit-exact lists, export bytes and FP32 phi against the FP32 oracle on the
same seeded inputs (both sides are built without FMA contraction and with
correctly rounded division and square root); FP32 phi within 1e-5 relative
(1e-5 absolute floor, |phi| <= 1) of an FP64 run of the same update; final
masks at Dice >= 0.9999 (in practice identical).
"""
import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import engine, synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Batch, Context, LssegError

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-5


def _say(msg):
    print(msg, flush=True)


def _assert_same_state(ctx: Context, ref: O.State, where=""):
    r, w, e, lst = ctx.state()
    rr, rw, re_ = ref.state()
    rl = ref.active()
    assert np.array_equal(lst, rl), f"{where}: lists differ (gpu {len(lst)} vs oracle {len(rl)})"
    assert np.array_equal(r.view(np.uint32), rr.view(np.uint32)), f"{where}: phi read differs"
    assert np.array_equal(w.view(np.uint32), rw.view(np.uint32)), f"{where}: phi write differs"
    assert np.array_equal(e, re_), f"{where}: export differs"


def _random_state(rng, N, n, wide=False):
    R = rng.uniform(-1, 1, N).astype(np.float32)
    if wide:
        R[rng.random(N) < 0.05] = 4.0
        R[rng.random(N) < 0.05] = -4.0
    # coarse levels give exact ties with the 0.008 / 0.1 thresholds
    lev = rng.random(N) < 0.3
    R[lev] = rng.choice(np.array([0.0, 0.008, 0.1, 0.5, -0.5, 1.0], np.float32), lev.sum())
    W = rng.uniform(-1, 1, N).astype(np.float32)
    E = rng.integers(-128, 128, N).astype(np.int8)
    V = rng.permutation(N)[:n].astype(np.uint32)
    return R, W, E, V


SOURCES = [
    (np.uint8, 0, 255, 150.0, 40.0), (np.int8, -128, 127, -10.0, 30.0),
    (np.uint16, 0, 4000, 1200.0, 250.0), (np.int16, -2000, 2000, 100.0, 300.0),
    (np.uint32, 0, 100000, 50000.0, 9000.0), (np.int32, -100000, 100000, -300.0, 20000.0),
    (np.float32, 0.0, 1.0, 0.55, 0.12),
]


@pytest.mark.parametrize("dtype,lo,hi,T,eps", SOURCES, ids=[np.dtype(s[0]).name for s in SOURCES])
@pytest.mark.parametrize("n", [2000, 6000], ids=["few", "many"])
def test_one_iteration_from_random_state(dtype, lo, hi, T, eps, n):
    _say(f"one iteration dtype={np.dtype(dtype).name} n={n}")
    rng = np.random.default_rng(n + sum(map(ord, np.dtype(dtype).name)))
    shape = (32, 32, 32)
    if np.issubdtype(dtype, np.floating):
        vol = rng.uniform(lo, hi, shape).astype(dtype)
    else:
        vol = rng.integers(lo, hi, shape).astype(dtype)
    ctx = Context(vol)
    ref = O.State(vol)
    R, W, E, V = _random_state(rng, ctx.N, n, wide=True)
    ctx.set_state(R, W, E, V)
    ref.set_state(R, W, E, V)
    assert O.path_is_few(n, ctx.N) == (n == 2000)
    for mode in (O.MODE_REF, O.MODE_CLEAN, O.MODE_REF):
        st = ctx.iterate(T, eps, 0.15, 0.5, 1, mode=mode)
        nh, ch, mh = ref.iterate(T, eps, 0.15, 0.5, 1, mode=mode)
        _assert_same_state(ctx, ref, f"mode {mode}")
        cnt = ctx.iteration_counts()
        assert (int(cnt.n[0]), int(cnt.c[0]), int(cnt.m[0])) == (int(nh[0]), int(ch[0]), int(mh[0]))
        assert st.sum_active == int(nh[0])
    ctx.close()


@pytest.mark.parametrize("dt", [0.5, -0.3])
def test_alpha_zero_and_negative_dt(dt):
    _say(f"alpha=0 dt={dt}")
    rng = np.random.default_rng(5)
    vol = rng.integers(0, 255, (16, 48, 32)).astype(np.uint8)
    ctx = Context(vol)
    ref = O.State(vol)
    R, W, E, V = _random_state(rng, ctx.N, 1500)
    ctx.set_state(R, W, E, V)
    ref.set_state(R, W, E, V)
    ctx.iterate(120, 50, 0.0, dt, 3)
    ref.iterate(120, 50, 0.0, dt, 3)
    _assert_same_state(ctx, ref)


def _run_pair(w, iters, chunk, mode=O.MODE_REF):
    ctx = Context(w.volume)
    ref = O.State(w.volume)
    ctx.set_phi(w.phi0, w.band)
    ref.set_phi(w.phi0, w.band)
    _assert_same_state(ctx, ref, "init")
    done = 0
    while done < iters:
        k = min(chunk, iters - done)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, k, mode=mode)
        nh, ch, mh = ref.iterate(w.T, w.eps, w.alpha, w.dt, k, mode=mode, threads=8)
        cnt = ctx.iteration_counts()
        assert np.array_equal(cnt.n, nh) and np.array_equal(cnt.c, ch) and np.array_equal(cnt.m, mh), \
            f"counts diverge in iterations {done}..{done + k}"
        done += k
        _assert_same_state(ctx, ref, f"after {done} iterations")
    return ctx, ref


def test_ellipsoids_small_multi_iteration():
    _say("ellipsoids 64^3, 120 iterations in chunks (graph and direct launches)")
    w = synth.ellipsoids(edge=64)
    ctx, ref = _run_pair(w, 120, 37)
    assert synth.dice(ctx.mask(), ref.mask()) == 1.0
    ctx.close()


def test_ellipsoids_full_workload():
    _say("ellipsoids 128^3, 300 iterations")
    w = synth.ellipsoids()
    ctx, ref = _run_pair(w, w.iterations, 100)
    m_gpu, m_ref = ctx.mask(), ref.mask()
    assert synth.dice(m_gpu, m_ref) >= 0.9999
    assert np.array_equal(m_gpu, m_ref)
    assert synth.dice(m_gpu, w.truth) > 0.9   # it segments the ellipsoids
    ctx.close()


def test_phantom_float_source_reduced():
    _say("phantom 181x217x181 float32, 150 iterations")
    w = synth.phantom()
    ctx, ref = _run_pair(w, 150, 75)
    assert np.array_equal(ctx.mask(), ref.mask())
    ctx.close()


def test_vessels_small_uint16():
    _say("vessels 128^3 uint16, 200 iterations")
    w = synth.vessels(edge=128, n_tubes=6)
    ctx, ref = _run_pair(w, 200, 100)
    assert np.array_equal(ctx.mask(), ref.mask())
    ctx.close()


def test_clean_mode_multi_iteration():
    _say("clean mode 48^3")
    w = synth.ellipsoids(edge=48)
    ctx, _ = _run_pair(w, 60, 30, mode=O.MODE_CLEAN)
    ctx.close()


def test_many_path_multi_iteration():
    _say("dense path: 32^3 with a wide band")
    n = 32
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    vol = ((x + y + z) % 7 * 30).astype(np.uint8)
    phi0 = np.clip((x - 15.5) / 3.0, -4, 4).astype(np.float32)
    w = synth.Workload("dense", vol, phi0, 100.0, 60.0, 0.2, 0.5, 40, band=3.5)
    assert not O.path_is_few(int((np.abs(phi0) < 3.5).sum()), n ** 3)
    ctx, ref = _run_pair(w, 40, 10)
    ctx2 = Context(vol)
    ctx2.set_phi(phi0, 3.5)
    st = ctx2.iterate(w.T, w.eps, w.alpha, w.dt, 40)
    assert st.many_path_iters > 0 and st.few_path_iters + st.many_path_iters == 40
    ctx.close()
    ctx2.close()


@pytest.mark.parametrize("wide", [False, True], ids=["few", "dense"])
def test_fused_and_separate_launches_agree(monkeypatch, wide):
    """Three launch schedules give identical states, counts and statistics:
    the default (each compaction inside the next update launch, which also
    rebuilds the traversal that the keep step left as kept directions and
    per-chunk counts), the same with a keep step that places the traversal
    itself through its look-back (LSSEG_DEFER_KEEP=0), and every step in a
    launch of its own (LSSEG_FUSE=0).  A profiled run of K iterations issues
    3 K + 1 launches, 3 K + 1 and 4 K."""
    _say(f"fused, placing-keep and separate launches, wide band {wide}")
    if wide:
        n = 32
        z, y, x = np.mgrid[0:n, 0:n, 0:n]
        vol = ((x + y + z) % 7 * 30).astype(np.uint8)
        phi0 = np.clip((x - 15.5) / 3.0, -4, 4).astype(np.float32)
        w = synth.Workload("dense", vol, phi0, 100.0, 60.0, 0.2, 0.5, 70, band=3.5)
    else:
        w = synth.ellipsoids(edge=64, iterations=70)
    out = {}
    for name, defer, fuse in (("deferred", "1", "1"), ("placing", "0", "1"), ("separate", "1", "0")):
        monkeypatch.setenv("LSSEG_DEFER_KEEP", defer)
        monkeypatch.setenv("LSSEG_FUSE", fuse)
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
        cnt = ctx.iteration_counts()
        state = ctx.state()
        ctx.set_phi(w.phi0, w.band)
        _, kt, per = ctx.iterate_profiled(w.T, w.eps, w.alpha, w.dt, 5)
        out[name] = (st, cnt, state, int(kt.launches), ctx.state())
        ctx.close()
    a = out["deferred"]
    for other in ("placing", "separate"):
        b = out[other]
        assert (a[0].few_path_iters, a[0].many_path_iters, a[0].sum_active, a[0].sum_candidates) == \
            (b[0].few_path_iters, b[0].many_path_iters, b[0].sum_active, b[0].sum_candidates), other
        for f in ("n", "c", "m"):
            assert np.array_equal(getattr(a[1], f), getattr(b[1], f)), (other, f)
        for x, y in zip(a[2] + a[4], b[2] + b[4]):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)), other
    if wide:
        assert a[0].many_path_iters > 0
    assert (out["deferred"][3], out["placing"][3], out["separate"][3]) == (3 * 5 + 1, 3 * 5 + 1, 4 * 5)


def test_calls_alternating_launch_schedules_match_the_oracle(monkeypatch):
    """One context whose calls alternate between the default schedule and
    separate launches (LSSEG_FUSE=0, read per call) ends where the oracle
    does.  A deferred call ends by rebuilding the traversal from the last keep
    step's kept directions, a separate one with the keep step placing it, and
    each starts from what the other left: every hand-over is exercised."""
    w = synth.ellipsoids(edge=64, iterations=0)
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    ref = O.State(w.volume)
    ref.set_phi(w.phi0, w.band)
    for i, (fuse, iters) in enumerate((("1", 9), ("0", 7), ("1", 12), ("0", 1), ("1", 1), ("1", 20), ("0", 5))):
        _say(f"call {i}: LSSEG_FUSE={fuse}, {iters} iterations")
        monkeypatch.setenv("LSSEG_FUSE", fuse)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, iters)
        ref.iterate(w.T, w.eps, w.alpha, w.dt, iters)
        _assert_same_state(ctx, ref, f"call {i}")
    ctx.close()


def test_direct_and_graph_launches_agree_past_the_stamp_wrap(monkeypatch):
    """600 iterations of the phantom, whose list stays non-empty, launched
    three ways: replayed graphs of 32 iterations, one launch at a time
    (LSSEG_GRAPH=0), and one launch at a time with per-launch events.  Each
    marking stamps the emission records, and the records are cleared and the
    stamp restarted twice in the call, between launches of either kind, with
    the previous compaction still to run.  Counts and states agree, and so
    does the batched form with and without graphs."""
    w = synth.phantom(iterations=600)
    out = {}
    for name in ("graphs", "direct", "profiled"):
        _say(f"phantom, 600 iterations, {name}")
        monkeypatch.setenv("LSSEG_GRAPH", "0" if name == "direct" else "1")
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        if name == "profiled":
            ctx.iterate_profiled(w.T, w.eps, w.alpha, w.dt, w.iterations)
        else:
            ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
        out[name] = (ctx.iteration_counts(), ctx.state())
        ctx.close()
    a = out["graphs"]
    assert int(np.count_nonzero(a[0].n)) == w.iterations, "the list should stay non-empty"
    for other in ("direct", "profiled"):
        b = out[other]
        for f in ("n", "c", "m"):
            assert np.array_equal(getattr(a[0], f), getattr(b[0], f)), (other, f)
        for x, y in zip(a[1], b[1]):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)), other
    prm = [(w.T, w.eps, w.alpha, w.dt, O.MODE_REF)]
    for graph in ("1", "0"):
        _say(f"phantom as a batch of one, 600 iterations, LSSEG_GRAPH={graph}")
        monkeypatch.setenv("LSSEG_GRAPH", graph)
        batch = Batch([w.volume])
        batch.volume(0).set_phi(w.phi0, w.band)
        batch.iterate(prm, w.iterations)
        cb = batch.volume(0).iteration_counts()
        assert np.array_equal(cb.n, a[0].n) and np.array_equal(cb.c, a[0].c), f"batch, graphs {graph}"
        r, wr, e, lst = batch.volume(0).state()
        assert np.array_equal(lst, a[1][3]) and np.array_equal(r.view(np.uint32), a[1][0].view(np.uint32))
        batch.close()


def test_many_calls_of_mixed_schedules_match_one_call(monkeypatch):
    """600 phantom iterations in 21 calls of uneven length (1 to 97
    iterations), each with its own schedule: graphs, direct launches
    (LSSEG_GRAPH=0) or separate launches (LSSEG_FUSE=0).  The stamp restarts
    land at every offset within a call and a graph chunk, and between calls;
    the end state equals one graph-replayed call of 600 iterations."""
    w = synth.phantom(iterations=600)
    ref = Context(w.volume)
    ref.set_phi(w.phi0, w.band)
    ref.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    want = ref.state()
    ref.close()
    sizes = [1, 7, 33, 97, 2, 31, 64, 5, 40, 1, 90, 3, 32, 17, 65, 11, 29, 8, 36, 20, 8]
    assert sum(sizes) == w.iterations
    schedules = [("1", "1"), ("0", "1"), ("1", "0")]
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    for i, k in enumerate(sizes):
        graph, fuse = schedules[i % 3]
        monkeypatch.setenv("LSSEG_GRAPH", graph)
        monkeypatch.setenv("LSSEG_FUSE", fuse)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, k)
    _say(f"{len(sizes)} calls done")
    for x, y in zip(ctx.state(), want):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    ctx.close()


@pytest.mark.parametrize("keep_grid,scan_grid", [("16", 16), ("24", 32)], ids=["keep-16", "keep-24-scan-32"])
def test_narrow_deferred_keep_grids_match_the_oracle(monkeypatch, keep_grid, scan_grid):
    """The deferred keep step on a grid narrow enough that each of its chunks
    holds many sub-tiles of 256 positions (LSSEG_KEEP_GRID with a small scan
    grid, which the keep grid never undercuts), and the update launch that
    rebuilds the traversal from those chunks: states equal the oracle's."""
    monkeypatch.setenv("LSSEG_KEEP_GRID", keep_grid)
    w = synth.ellipsoids(edge=64, iterations=80)
    _say(f"keep grid {keep_grid}, scan grid {scan_grid}, {w.iterations} iterations")
    ctx = Context(w.volume)
    ctx.set_tuning(0, scan_grid)
    ctx.set_phi(w.phi0, w.band)
    ref = O.State(w.volume)
    ref.set_phi(w.phi0, w.band)
    for leg in (33, 47):
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, leg)
        ref.iterate(w.T, w.eps, w.alpha, w.dt, leg)
        _assert_same_state(ctx, ref, f"keep grid {keep_grid} after {leg}")
    assert int(ctx.iteration_counts().n.max()) > 16 * 256 * 2, "chunks should hold several sub-tiles"
    ctx.close()


@pytest.mark.parametrize("pack", ["1", "0"], ids=["packed-counts", "one-count-per-word"])
def test_compaction_tile_sizes_agree(monkeypatch, pack):
    """Few-path compactions take tiles of 1024 list positions, or of 256 for
    lists up to LSSEG_SMALL_TILE_N positions, and the workgroups that load
    their tile together with the counts keep whichever layout the plan needs.
    Large tiles only, small tiles only, and a threshold inside the run's range
    of list sizes give the same states, counts and statistics.  With
    LSSEG_LOOKBACK_PACK=0 each chunk publishes its seven counts one to a
    look-back word (the layout of chunks of 2^16 positions or more)."""
    monkeypatch.setenv("LSSEG_LOOKBACK_PACK", pack)
    w = synth.ellipsoids(edge=64, iterations=70)
    out = {}
    for small in ("0", "1000000000", None):
        if small is None:  # the median list size of the first run: the run crosses it
            small = str(int(np.median(out["0"][1].n)))
        _say(f"compaction tiles: small up to {small} positions, LSSEG_LOOKBACK_PACK={pack}")
        monkeypatch.setenv("LSSEG_SMALL_TILE_N", small)
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
        out[small] = (st, ctx.iteration_counts(), ctx.state())
        ctx.close()
    runs = list(out.values())
    n = runs[0][1].n
    assert n.min() <= int(list(out)[2]) < n.max()
    for b in runs[1:]:
        assert runs[0][0].sum_active == b[0].sum_active
        for f in ("n", "c", "m"):
            assert np.array_equal(getattr(runs[0][1], f), getattr(b[1], f)), f
        for x, y in zip(runs[0][2], b[2]):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


@pytest.mark.parametrize("pack", ["1", "0"], ids=["packed-counts", "one-count-per-word"])
@pytest.mark.parametrize("wide", [False, True], ids=["few", "dense"])
def test_chunks_of_several_subtiles(monkeypatch, wide, pack):
    """With a 16-workgroup grid the keep step, the compaction (8 workgroups in
    the fused launch) and the dense initial compaction each take several
    sub-tiles per look-back chunk (counted first, placed after the prefix):
    states and counts still equal the oracle's every chunk.  The few-path
    compaction's chunks publish their seven counts two to a look-back word, or
    (LSSEG_LOOKBACK_PACK=0) one to a word."""
    _say(f"16-workgroup grid, wide band {wide}, LSSEG_LOOKBACK_PACK={pack}")
    monkeypatch.setenv("LSSEG_GRID_BLOCKS", "16")
    monkeypatch.setenv("LSSEG_LOOKBACK_PACK", pack)
    if wide:
        n = 32
        z, y, x = np.mgrid[0:n, 0:n, 0:n]
        vol = ((x + y + z) % 7 * 30).astype(np.uint8)
        phi0 = np.clip((x - 15.5) / 3.0, -4, 4).astype(np.float32)
        w = synth.Workload("dense", vol, phi0, 100.0, 60.0, 0.2, 0.5, 40, band=3.5)
        iters, chunk, need = 40, 13, 8 * 2048 // 7  # several sub-tiles per compaction chunk
    else:
        w = synth.ellipsoids(edge=64)
        iters, chunk, need = 120, 41, 16 * 512  # several sub-tiles per keep chunk
    probe = O.State(w.volume)
    probe.set_phi(w.phi0, w.band)
    assert max(probe.iterate(w.T, w.eps, w.alpha, w.dt, iters, threads=8)[0]) > need
    ctx, ref = _run_pair(w, iters, chunk)
    assert np.array_equal(ctx.mask(), ref.mask())
    ctx.close()


def test_empty_list_ends_the_run_early(monkeypatch):
    """vessels 64^3 with 4 tubes empties its list at iteration 101; of 400
    iterations the trailing ones over the empty list are not launched.  Counts,
    statistics, both phi buffers (and which one is current), export and mask
    equal the oracle's full run and a run that launches every iteration."""
    _say("vessels 64^3, 400 iterations, list empty from iteration 101")
    monkeypatch.setenv("LSSEG_GRAPH", "1")  # the early stop watches the graph-replayed chunks
    w = synth.vessels(edge=64, n_tubes=4)
    ref = O.State(w.volume)
    ref.set_phi(w.phi0, w.band)
    nh, ch, mh = ref.iterate(w.T, w.eps, w.alpha, w.dt, 400, threads=8)
    assert nh[-1] == 0 and int(np.argmax(np.asarray(nh) == 0)) < 150
    runs = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("LSSEG_FUSE", fuse)
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, 400)
        cnt = ctx.iteration_counts()
        assert np.array_equal(cnt.n, nh) and np.array_equal(cnt.c, ch) and np.array_equal(cnt.m, mh), fuse
        _assert_same_state(ctx, ref, f"LSSEG_FUSE={fuse}")
        assert np.array_equal(ctx.phi(), ref.phi()), fuse
        assert np.array_equal(ctx.mask(), ref.mask()), fuse
        runs[fuse] = st
        # iterating on from the empty list is still exact
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, 3)
        ctx.close()
    a, b = runs["1"], runs["0"]
    assert a.empty_skipped > 150 and b.empty_skipped == 0
    many = sum(1 for n in nh if not O.path_is_few(int(n), ctx.N))
    assert (a.few_path_iters, a.many_path_iters, a.sum_active, a.active_final) == \
        (b.few_path_iters, b.many_path_iters, b.sum_active, b.active_final) == (400 - many, many, sum(nh), 0)


def test_fp64_tolerance_per_iteration():
    _say("fp64 tolerance along the ellipsoids run")
    w = synth.ellipsoids(edge=64)
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    X, Y, Z = ctx.X, ctx.Y, ctx.Z
    src = O.tile_volume(np.pad(w.volume, [(0, Z - w.volume.shape[0]), (0, Y - w.volume.shape[1]),
                                          (0, X - w.volume.shape[2])]))
    worst = 0.0
    for step in range(0, 90, 3):
        r, _, _, lst = ctx.state()
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, 1)
        new_r = ctx.state()[0]
        W64, _, _ = O.update(X, Y, Z, src, r.astype(np.float64), w.T, w.eps, w.alpha, w.dt, lst)
        got = new_r[lst].astype(np.float64)
        want = W64[lst]
        err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
        worst = max(worst, float(err.max()) if len(lst) else 0.0)
        ctx.iterate(w.T, w.eps, w.alpha, w.dt, 2)
    _say(f"worst relative error vs fp64: {worst:.3e}")
    assert worst <= FP64_RTOL
    ctx.close()


def test_seed_sphere_and_iterate():
    _say("seed sphere 40x36x33 (padded), anisotropy 1.5")
    rng = np.random.default_rng(11)
    vol = np.clip(rng.normal(100, 15, (33, 36, 40)), 0, 255).astype(np.uint8)
    vol[8:25, 10:28, 10:30] = np.clip(rng.normal(200, 10, (17, 18, 20)), 0, 255).astype(np.uint8)
    ctx = Context(vol, z_aniso=1.5)
    ref = O.State(vol, z_aniso=1.5)
    for c, r in (((20.6, 19.2, 16.9), 5), ((12.0, 14.0, 12.0), 3)):   # second seed unions
        ctx.seed_sphere(*c, r)
        ref.seed_sphere(*c, r)
        _assert_same_state(ctx, ref, f"seed {c}")
        ctx.iterate(200, 50, 0.2, 0.4, 25)
        ref.iterate(200, 50, 0.2, 0.4, 25)
        _assert_same_state(ctx, ref, f"after seed {c}")
    ctx.close()


def test_reinit_freeze_clear_and_mask():
    _say("parameter-change restart, freeze, clear")
    w = synth.ellipsoids(edge=48)
    ctx = Context(w.volume)
    ref = O.State(w.volume)
    for s in (ctx, ref):
        s.set_phi(w.phi0, 1.0)
    ctx.iterate(w.T, w.eps, w.alpha, w.dt, 20)
    ref.iterate(w.T, w.eps, w.alpha, w.dt, 20)
    ctx.iterate_reinit(w.T - 10, w.eps, w.alpha, w.dt)
    ref.iterate_reinit(w.T - 10, w.eps, w.alpha, w.dt)
    _assert_same_state(ctx, ref, "reinit")
    ctx.iterate(w.T - 10, w.eps, w.alpha, w.dt, 15)
    ref.iterate(w.T - 10, w.eps, w.alpha, w.dt, 15)
    _assert_same_state(ctx, ref, "after reinit")
    assert np.array_equal(ctx.mask(), ref.mask())
    ctx.freeze()
    ref.freeze()
    ctx.clear(False)
    ref.clear(False)
    assert np.array_equal(ctx.mask(), ref.mask()) and ctx.mask().sum() > 0
    assert ctx.active_count == 0
    ctx.clear(True)
    assert ctx.mask().sum() == 0
    ctx.close()


def test_phi_readback_unpadded():
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, (19, 21, 37)).astype(np.float32)
    phi0 = synth.sphere_sdf(37, 21, 19, (18, 10, 9), 4.0)
    w = synth.Workload("odd", vol, phi0, 0.5, 0.3, 0.1, 0.5, 5)
    ctx = Context(w.volume)
    ref = O.State(w.volume)
    ctx.set_phi(w.phi0, 1.0)
    ref.set_phi(w.phi0, 1.0)
    ctx.iterate(w.T, w.eps, w.alpha, w.dt, 5)
    ref.iterate(w.T, w.eps, w.alpha, w.dt, 5)
    for which in (0, 1):
        assert np.array_equal(ctx.phi(which), ref.phi(which))
    ctx.close()


def test_errors_are_reported_not_aborted():
    vol = np.zeros((16, 16, 16), np.uint8)
    ctx = Context(vol)
    with pytest.raises(LssegError) as e:
        ctx.iterate(100, 0.5, 0.1, 0.5, 1)      # integer sources need eps >= 1
    assert e.value.status == engine.ERR_INVALID_ARG
    with pytest.raises(LssegError):
        ctx.iterate(100, 10, float("nan"), 0.5, 1)
    with pytest.raises(LssegError):
        ctx.iterate(100, 10, 0.1, 0.5, 1, mode=7)
    with pytest.raises(LssegError) as e:
        ctx.set_state(np.zeros(ctx.N, np.float32), np.zeros(ctx.N, np.float32), np.zeros(ctx.N, np.int8),
                      np.array([1, 1], np.uint32))
    assert "duplicate" in str(e.value)
    with pytest.raises(LssegError):
        ctx.set_state(np.zeros(ctx.N, np.float32), np.zeros(ctx.N, np.float32), np.zeros(ctx.N, np.int8),
                      np.array([ctx.N], np.uint32))
    # the context still works after rejected calls
    ctx.set_phi(np.full((16, 16, 16), 0.5, np.float32), 1.0)
    assert ctx.active_count == 4096
    ctx.iterate(100, 10, 0.1, 0.5, 2)
    ctx.close()
