"""The batched form against solo runs and the oracle, through the C ABI (needs a GPU).

Bar: every volume of a batch ends bit-identical to a solo ``lsseg_iterate`` run
of the same volume with the same parameters (active list, both phi buffers,
export bytes, mask, per-iteration counts), and to the oracle.  The batch mixes
uneven shapes, per-volume parameters, a volume whose list empties, one on the
dense ("many") path, and volumes entering with either buffer parity; the
iteration counts cross the captured-graph chunk size.
"""
import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Batch, Context, LssegError

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _volumes():
    """(volume, phi0, band, params) for six uneven volumes."""
    out = []
    shapes = [(20, 33, 47), (64, 40, 24), (31, 31, 31), (17, 70, 25), (48, 48, 48), (24, 24, 24)]
    for i, (nz, ny, nx) in enumerate(shapes):
        vol, _ = synth.ellipsoid_volume(nx, ny, nz, synth.SEED_BASE + 77 + i,
                                     [(nx / 2.0, ny / 2.0, nz / 2.0, nx / 3.0, ny / 3.5, nz / 3.2)],
                                     sigma=8.0)
        phi0 = synth.sphere_sdf(nx, ny, nz, (nx / 2.0, ny / 2.0, nz / 2.0), max(2.0, min(nx, ny, nz) / 6.0))
        band = 1.0
        T = float(150 + 10 * i)
        eps = float(25 + 3 * i)
        alpha = 0.05 + 0.03 * i
        mode = O.MODE_REF if i != 3 else O.MODE_CLEAN
        if i == 2:
            band = 3.5  # wide band on a small volume: the dense path
        if i == 5:
            T, eps = 250.0, 2.0  # nothing inside the window: the list empties
        out.append((vol, phi0, band, (T, eps, alpha, 0.5, mode)))
    return out


def _assert_same(a: Context, b: Context, where):
    ra, wa, ea, la = a.state()
    rb, wb, eb, lb = b.state()
    assert np.array_equal(la, lb), f"{where}: lists differ ({len(la)} vs {len(lb)})"
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32)), f"{where}: phi read differs"
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"{where}: phi write differs"
    assert np.array_equal(ea, eb), f"{where}: export differs"
    assert np.array_equal(a.mask(), b.mask()), f"{where}: mask differs"


def test_batch_matches_solo_runs_and_oracle():
    specs = _volumes()
    _say(f"batch of {len(specs)} volumes")
    batch = Batch([s[0] for s in specs])
    solos = [Context(s[0]) for s in specs]
    for i, (vol, phi0, band, prm) in enumerate(specs):
        batch.volume(i).set_phi(phi0, band)
        solos[i].set_phi(phi0, band)
    # volumes 1 and 4 enter the batch with the other buffer parity
    for i in (1, 4):
        T, eps, alpha, dt, mode = specs[i][3]
        batch.volume(i).iterate(T, eps, alpha, dt, 1, mode=mode)
        solos[i].iterate(T, eps, alpha, dt, 1, mode=mode)
    params = [s[3] for s in specs]
    many = [0] * len(specs)
    for leg, iters in enumerate((70, 37)):
        _say(f"leg {leg}: {iters} iterations")
        stats = batch.iterate(params, iters)
        many = [a + int(st.many_path_iters) for a, st in zip(many, stats)]
        for i, s in enumerate(solos):
            T, eps, alpha, dt, mode = params[i]
            st = s.iterate(T, eps, alpha, dt, iters, mode=mode)
            assert stats[i].sum_active == st.sum_active, f"volume {i}"
            assert stats[i].sum_candidates == st.sum_candidates, f"volume {i}"
            assert stats[i].active_final == st.active_final, f"volume {i}"
            assert stats[i].many_path_iters == st.many_path_iters, f"volume {i}"
            cb, cs = batch.volume(i).iteration_counts(), s.iteration_counts()
            assert np.array_equal(cb.n, cs.n) and np.array_equal(cb.c, cs.c) and np.array_equal(cb.m, cs.m)
            _assert_same(batch.volume(i), s, f"leg {leg} volume {i}")
    assert many[2] > 0, "volume 2 should take the dense path"
    assert stats[5].active_final == 0, "volume 5 should empty its list"
    # volume 0 against the oracle from the same start
    vol, phi0, band, (T, eps, alpha, dt, mode) = specs[0]
    ref = O.State(vol)
    ref.set_phi(phi0, band)
    ref.iterate(T, eps, alpha, dt, 107, mode=mode)
    assert np.array_equal(batch.volume(0).active(), ref.active())
    assert np.array_equal(batch.volume(0).mask(), ref.mask())
    for s in solos:
        s.close()
    batch.close()


def test_batch_with_the_smallest_scan_grid():
    """One more scan workgroup than volumes: the batched keep and compaction
    take several sub-tiles per look-back chunk; every volume still equals its
    solo run (default grid)."""
    specs = _volumes()
    _say(f"batch of {len(specs)} volumes, {len(specs) + 1} scan workgroups")
    batch = Batch([s[0] for s in specs])
    batch.set_tuning(256, len(specs) + 1)
    solos = [Context(s[0]) for s in specs]
    for i, (vol, phi0, band, prm) in enumerate(specs):
        batch.volume(i).set_phi(phi0, band)
        solos[i].set_phi(phi0, band)
    params = [s[3] for s in specs]
    batch.iterate(params, 45)
    for i, s in enumerate(solos):
        T, eps, alpha, dt, mode = params[i]
        s.iterate(T, eps, alpha, dt, 45, mode=mode)
        cb, cs = batch.volume(i).iteration_counts(), s.iteration_counts()
        assert np.array_equal(cb.n, cs.n) and np.array_equal(cb.c, cs.c) and np.array_equal(cb.m, cs.m)
        _assert_same(batch.volume(i), s, f"volume {i}")
    with pytest.raises(LssegError):
        batch.set_tuning(256, len(specs))  # not above the volume count
    for s in solos:
        s.close()
    batch.close()


def test_batch_checkpoint_and_autotune_keep_results():
    specs = _volumes()[:4]
    batch = Batch([s[0] for s in specs])
    for i, (vol, phi0, band, prm) in enumerate(specs):
        batch.volume(i).set_phi(phi0, band)
    params = [s[3] for s in specs]
    batch.checkpoint_save(0)
    batch.iterate(params, 40)
    first = [batch.volume(i).state() for i in range(len(specs))]
    t = batch.autotune(params, 20)
    _say(f"autotune: {t.update_blocks} update blocks, {t.compact_blocks} compaction blocks, keep items "
         f"{t.keep_items}, compaction tile {t.compact_tile}, {t.candidates} tunings, {t.ms:.3f} ms")
    assert t.candidates >= 4 and t.update_blocks > 0 and t.compact_blocks > 0 and not t.cached
    assert t.keep_items in (1, 2) and t.compact_tile in (256, 1024)
    got = batch.tuning()
    assert (got.update_blocks, got.keep_items, got.compact_tile) == (t.update_blocks, t.keep_items, t.compact_tile)
    # the same workload again: the cached result, applied without timing
    t2 = batch.autotune(params, 20)
    assert t2.cached and t2.candidates == 0
    assert (t2.update_blocks, t2.compact_blocks, t2.keep_items, t2.compact_tile) == \
        (t.update_blocks, t.compact_blocks, t.keep_items, t.compact_tile)
    batch.checkpoint_restore(0)
    batch.iterate(params, 40)
    for i in range(len(specs)):
        r, w, e, lst = batch.volume(i).state()
        assert np.array_equal(lst, first[i][3]) and np.array_equal(r.view(np.uint32), first[i][0].view(np.uint32))
    with pytest.raises(LssegError):
        batch.set_tuning(0, 1)
    with pytest.raises(LssegError):
        batch.set_tuning(256, 64, keep_items=3)
    batch.close()


@pytest.mark.parametrize("keep_items,compact_tile", [(1, 256), (2, 256), (1, 1024)])
def test_batch_widths_give_the_solo_results(keep_items, compact_tile):
    """Every tuning of the keep step's width and the compaction tile leaves each
    volume bit-identical to its solo run."""
    specs = _volumes()
    batch = Batch([s[0] for s in specs])
    t = batch.tuning()
    batch.set_tuning(t.update_blocks, t.compact_blocks, keep_items, compact_tile)
    solos = [Context(s[0]) for s in specs]
    for i, (vol, phi0, band, prm) in enumerate(specs):
        batch.volume(i).set_phi(phi0, band)
        solos[i].set_phi(phi0, band)
    params = [s[3] for s in specs]
    batch.iterate(params, 41)
    for i, s in enumerate(solos):
        T, eps, alpha, dt, mode = params[i]
        s.iterate(T, eps, alpha, dt, 41, mode=mode)
        _assert_same(batch.volume(i), s, f"widths {keep_items}/{compact_tile} volume {i}")
        s.close()
    batch.close()


def test_batch_rejects_mixed_types_and_owned_destroy():
    a = np.zeros((16, 16, 16), np.uint8)
    b = np.zeros((16, 16, 16), np.uint16)
    with pytest.raises(LssegError):
        Batch([a, b])
    batch = Batch([a, a])
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import lib, ERR_STATE
    assert lib().lsseg_destroy(batch.volume(0)._h) == ERR_STATE
    batch.close()


@pytest.mark.parametrize("defer,fuse,keep_grid", [("0", "1", None), ("1", "0", None), ("1", "1", "64")],
                         ids=["placing-keep", "separate-launches", "narrow-keep-grid"])
def test_batch_schedules_give_the_solo_results(monkeypatch, defer, fuse, keep_grid):
    """The batch's other schedules leave each volume bit-identical to its solo
    run: the keep step that places the traversals itself (LSSEG_DEFER_KEEP=0),
    every step in a launch of its own (LSSEG_FUSE=0), and the deferred keep
    step on a grid narrow enough that its chunks hold several sub-tiles."""
    specs = _volumes()
    monkeypatch.setenv("LSSEG_DEFER_KEEP", defer)
    monkeypatch.setenv("LSSEG_FUSE", fuse)
    if keep_grid:
        monkeypatch.setenv("LSSEG_BATCH_KEEP_GRID", keep_grid)
    batch = Batch([s[0] for s in specs])
    t = batch.tuning()
    if keep_grid:
        batch.set_tuning(t.update_blocks, 64, t.keep_items, t.compact_tile)
    monkeypatch.setenv("LSSEG_DEFER_KEEP", "1")
    monkeypatch.setenv("LSSEG_FUSE", "1")
    solos = [Context(s[0]) for s in specs]
    for i, (vol, phi0, band, prm) in enumerate(specs):
        batch.volume(i).set_phi(phi0, band)
        solos[i].set_phi(phi0, band)
    params = [s[3] for s in specs]
    monkeypatch.setenv("LSSEG_FUSE", fuse)
    batch.iterate(params, 41)
    monkeypatch.setenv("LSSEG_FUSE", "1")
    for i, s in enumerate(solos):
        T, eps, alpha, dt, mode = params[i]
        s.iterate(T, eps, alpha, dt, 41, mode=mode)
        _assert_same(batch.volume(i), s, f"schedule defer={defer} fuse={fuse} keep grid {keep_grid} volume {i}")
        s.close()
    batch.close()
