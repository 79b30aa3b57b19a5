"""Fixed-point phi storage (LSSEG_PHI_I8, the reference's LEVEL_SET_CHAR) on the
GPU against the oracle's fixed-point mode, through the C ABI (needs a GPU).

Bar: bit-exact states (the stored values, read back as s / 127), export bytes
and active lists, from random states on both compaction paths, over full runs
from phi0 and from the seed sphere, and per volume in a batch.
"""
import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import PHI_I8, Batch, Context, LssegError

pytestmark = pytest.mark.gpu


def _same(ctx, ref, where):
    r, w, e, lst = ctx.state()
    rr, rw, re_ = ref.state()
    assert np.array_equal(lst, ref.active()), f"{where}: lists"
    assert np.array_equal(r.view(np.uint32), rr.view(np.uint32)), f"{where}: phi read"
    assert np.array_equal(w.view(np.uint32), rw.view(np.uint32)), f"{where}: phi write"
    assert np.array_equal(e, re_), f"{where}: export"


@pytest.mark.parametrize("n", [1500, 6000], ids=["few", "many"])
def test_random_state_iterations(n):
    print(f"int8 random state n={n}", flush=True)
    rng = np.random.default_rng(31 + n)
    vol = rng.integers(0, 255, (32, 32, 32)).astype(np.uint8)
    ctx = Context(vol, phi_storage=PHI_I8)
    ref = O.State(vol, phi_storage=1)
    k = rng.integers(-128, 128, ctx.N)
    R = (k.astype(np.float32) / np.float32(127.0)).astype(np.float32)
    R[k == -128] = -1.0
    W = R[rng.permutation(ctx.N)]
    E = rng.integers(-128, 128, ctx.N).astype(np.int8)
    V = rng.permutation(ctx.N)[:n].astype(np.uint32)
    ctx.set_state(R, W, E, V)
    ref.set_state(R, W, E, V)
    _same(ctx, ref, "after set_state")
    for it in range(4):
        ctx.iterate(150.0, 60.0, 0.2, 0.5, 1)
        ref.iterate(150.0, 60.0, 0.2, 0.5, 1)
        _same(ctx, ref, f"iteration {it}")
    ctx.close()


def test_full_runs_from_phi0_and_seed():
    w = synth.ellipsoids(edge=64, iterations=150)
    print("int8 ellipsoids 64^3, 150 iterations", flush=True)
    ctx = Context(w.volume, phi_storage=PHI_I8)
    ref = O.State(w.volume, phi_storage=1)
    ctx.set_phi(w.phi0, w.band)
    ref.set_phi(w.phi0, w.band)
    _same(ctx, ref, "set_phi")
    ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    ref.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    _same(ctx, ref, "after the run")
    assert np.array_equal(ctx.mask(), ref.mask())
    # seed sphere and a parameter-change restart
    ctx.clear(True)
    ref.clear(1)
    ctx.seed_sphere(*w.seed_center, 5)
    ref.seed_sphere(*w.seed_center, 5)
    _same(ctx, ref, "seed")
    ctx.iterate(w.T, w.eps, w.alpha, w.dt, 30)
    ref.iterate(w.T, w.eps, w.alpha, w.dt, 30)
    ctx.iterate_reinit(w.T - 5, w.eps, w.alpha, w.dt)
    ref.iterate_reinit(w.T - 5, w.eps, w.alpha, w.dt)
    _same(ctx, ref, "reinit")
    with pytest.raises(LssegError):
        ctx.fwd_record(w.T, w.eps, w.alpha, w.dt, 5)
    ctx.close()


def test_batch_of_fixed_point_volumes():
    ws = [synth.ellipsoids(edge=e, iterations=90) for e in (32, 48)]
    b = Batch([w.volume for w in ws], phi_storage=PHI_I8)
    refs = []
    for i, w in enumerate(ws):
        b.volume(i).set_phi(w.phi0, w.band)
        r = O.State(w.volume, phi_storage=1)
        r.set_phi(w.phi0, w.band)
        refs.append(r)
    b.iterate([(w.T, w.eps, w.alpha, w.dt, 0) for w in ws], 90)
    for i, (w, r) in enumerate(zip(ws, refs)):
        r.iterate(w.T, w.eps, w.alpha, w.dt, 90)
        _same(b.volume(i), r, f"volume {i}")
    b.close()
