"""The differentiable form on the GPU, through the C ABI (needs a GPU).

Bar:
  - the recording run's lists and decision bytes equal the oracle's FP32
    recording bit for bit, iteration by iteration as sets of (voxel, decision)
    pairs (the engine records each V_k in its traversal order, the oracle in
    the reference's list order), and the final state equals a plain run's;
  - the replay at the recorded inputs reproduces the forward phi bit for bit;
  - the GPU adjoint (FP32, hand-written kernel) agrees with the oracle's FP64
    adjoint of the same recorded map: dL/dT, dL/deps, dL/dalpha within 1e-3
    relative and dL/dphi0 within 1e-3 relative (L2);
  - it agrees with central finite differences (h = 1e-3, five-point stencil)
    of the GPU replay to 1e-3 relative, and the torch.autograd.Function gives
    the same gradients.  (The three-point stencil's truncation error at
    h = 1e-3 is itself ~1e-3 relative on this 40-iteration map in FP64; the
    five-point stencil's is ~5e-5.)
"""
import numpy as np
import pytest

# torch ships its own HIP runtime and must provide the process's runtime: it is
# imported here, at collection, before any test loads the engine.
torch = pytest.importorskip("torch")

import oracle as O  # noqa: E402
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth  # noqa: E402
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context  # noqa: E402

pytestmark = pytest.mark.gpu

DT = 0.5
RTOL = 1e-3


def _say(msg):
    print(msg, flush=True)


def _case(edge=(28, 24, 20), seed=12345, iters=25):
    nx, ny, nz = edge
    vol, gt = synth.ellipsoid_volume(nx, ny, nz, seed, [(nx / 2, ny / 2, nz / 2, nx / 3, ny / 3.4, nz / 3.3)],
                                     as_float_scale=1 / 255.0, truth=True)
    phi0 = synth.sphere_sdf(nx, ny, nz, (nx / 2, ny / 2, nz / 2), max(2.0, min(edge) / 7.0))
    return vol, gt, phi0, (0.6, 0.15, 0.1), iters


def _soft_dice_and_grad(phi, mask, temp=0.5):
    """loss and dloss/dphi in float64 (host)."""
    phi = phi.astype(np.float64)
    m = mask.astype(np.float64)
    s = 1.0 / (1.0 + np.exp(phi / temp))
    a = (s * m).sum()
    b = s.sum() + m.sum()
    loss = 1.0 - 2.0 * a / b
    dl_ds = -2.0 * (m * b - a) / (b * b)
    ds_dphi = -s * (1.0 - s) / temp
    return loss, dl_ds * ds_dphi


def _same_record(n, lists, dec, on, olists, odec):
    """V_k and its decisions against the oracle's, per iteration, after sorting
    both by voxel (the voxels of one V_k are distinct)."""
    if not np.array_equal(n, on) or lists.size != olists.size:
        return False
    off = 0
    for k in n.astype(np.int64):
        sl = slice(off, off + int(k))
        i = np.argsort(lists[sl], kind="stable")
        j = np.argsort(olists[sl], kind="stable")
        if not (np.array_equal(lists[sl][i], olists[sl][j]) and np.array_equal(dec[sl][i], odec[sl][j])):
            return False
        off += int(k)
    return True


def test_record_and_replay_match_oracle_and_forward():
    vol, _, phi0, (T, eps, alpha), K = _case()
    _say(f"record {vol.shape} K={K}")
    ctx = Context(vol)
    ctx.set_phi(phi0, 1.0)
    rec = ctx.fwd_record(T, eps, alpha, DT, K)
    n, lists, dec = rec.lists()
    on, olists, odec, ost = O.record_f32(vol, phi0, 1.0, T, eps, alpha, DT, K)
    assert _same_record(n, lists, dec, on, olists, odec)
    plain = Context(vol)
    plain.set_phi(phi0, 1.0)
    plain.iterate(T, eps, alpha, DT, K)
    for a, b in zip(ctx.state(), plain.state()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    out = rec.replay(phi0, T, eps, alpha, DT)
    assert np.array_equal(out.view(np.uint32), ctx.phi(0).view(np.uint32)), "replay differs from the forward"
    rec.close()
    ctx.close()
    plain.close()


@pytest.mark.parametrize("capacity", ["1", ""], ids=["short-first-guess", "guess-from-last-record"])
def test_record_sizes_itself(monkeypatch, capacity):
    """One recording run sizes the record from device offsets.  With a first
    guess that is too short (forced to one entry) the record runs again from
    the start with the exact size; without it, the second record of a context
    starts from the first one's size.  Either way the lists and decisions are
    the oracle's and the context ends in the plain run's state.  The backward
    pass steps back from the record's own copy of the final buffers, so a
    second call on the same record, and the other record, give the same
    gradients (up to the order of the adjoint's float atomics)."""
    vol, _, phi0, (T, eps, alpha), K = _case(edge=(30, 26, 22), seed=777, iters=30)
    if capacity:
        monkeypatch.setenv("LSSEG_RECORD_CAPACITY", capacity)
    _say(f"record sizing {vol.shape} K={K} capacity={capacity or 'auto'}")
    on, olists, odec, _ = O.record_f32(vol, phi0, 1.0, T, eps, alpha, DT, K)
    plain = Context(vol)
    plain.set_phi(phi0, 1.0)
    plain.iterate(T, eps, alpha, DT, K)
    ctx = Context(vol)
    dphi = np.random.default_rng(5).standard_normal(vol.shape).astype(np.float32)
    grads = []
    for rep in range(2):
        ctx.set_phi(phi0, 1.0)
        rec = ctx.fwd_record(T, eps, alpha, DT, K)
        n, lists, dec = rec.lists()
        assert _same_record(n, lists, dec, on, olists, odec), rep
        for a, b in zip(ctx.state(), plain.state()):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), rep
        out = rec.replay(phi0, T, eps, alpha, DT)
        assert np.array_equal(out.view(np.uint32), ctx.phi(0).view(np.uint32)), rep
        g0, th = rec.bwd(dphi)
        g0b, thb = rec.bwd(dphi)
        assert np.allclose(g0b, g0, rtol=1e-5, atol=1e-6) and np.allclose(thb, th, rtol=1e-5), rep
        grads.append((g0, th))
        rec.close()
    assert np.abs(grads[0][0]).max() > 0
    assert np.allclose(grads[0][0], grads[1][0], rtol=1e-5, atol=1e-6)
    assert np.allclose(grads[0][1], grads[1][1], rtol=1e-5)
    ctx.close()
    plain.close()


def test_adjoint_matches_fp64_oracle_and_finite_differences():
    vol, gt, phi0, (T, eps, alpha), K = _case(edge=(40, 36, 30), iters=40)
    nz, ny, nx = vol.shape
    _say(f"adjoint {vol.shape} K={K}")
    ctx = Context(vol)
    ctx.set_phi(phi0, 1.0)
    rec = ctx.fwd_record(T, eps, alpha, DT, K)
    phiK = ctx.phi(0)
    loss, dphi = _soft_dice_and_grad(phiK, gt)
    g0, th = rec.bwd(dphi.astype(np.float32))
    # FP64 host adjoint of the same recorded map
    n, lists, dec = rec.lists()
    X, Y, Z = O.padded_dims(nx, ny, nz)
    src = O.padded_tiled_source(vol)
    p0 = O.pad_tile_phi(phi0, 1.0)
    out64, hist = O.replay_f64(X, Y, Z, src, T, eps, alpha, DT, n, lists, dec, p0, keep_history=True)
    phiK64 = O.unpad_untile(out64, nx, ny, nz)
    assert np.abs(phiK64 - phiK).max() < 1e-4
    _, dphi64 = _soft_dice_and_grad(phiK64, gt)
    g64, th64 = O.adjoint_f64(X, Y, Z, src, T, eps, alpha, DT, n, lists, dec, hist, O.pad_tile_phi(dphi64, 0.0))
    g64 = O.unpad_untile(g64, nx, ny, nz)
    _say(f"dtheta gpu {th} fp64 {th64}")
    for i in range(3):
        assert abs(th[i] - th64[i]) <= RTOL * abs(th64[i]), (i, th[i], th64[i])
    assert np.linalg.norm(g0 - g64) <= RTOL * np.linalg.norm(g64)
    # central finite differences of the GPU replay (h = 1e-3, five-point stencil)
    h = 1e-3

    def L(p, t, e, a):
        return _soft_dice_and_grad(rec.replay(p, t, e, a, DT), gt)[0]

    def central(f):
        return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)

    base = [T, eps, alpha]
    for i in range(3):
        def f(step, i=i):
            v = list(base)
            v[i] += step
            return L(phi0, *v)
        fd = central(f)
        _say(f"param {i}: adjoint {th[i]:.6e} fd {fd:.6e}")
        assert abs(th[i] - fd) <= RTOL * max(abs(fd), 1e-3), (i, th[i], fd)
    # along the gradient (the derivative there is |dL/dphi0|, well above rounding)
    # mixed with a random component near the front
    rng = np.random.default_rng(3)
    r = rng.standard_normal(phi0.shape) * (np.abs(phi0) < 3.0)
    d = g0 / np.linalg.norm(g0) + 0.5 * r / np.linalg.norm(r)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    fd = central(lambda step: L(phi0 + step * d, *base))
    adj = float((g0.astype(np.float64) * d).sum())
    _say(f"phi0 direction: adjoint {adj:.6e} fd {fd:.6e}")
    assert abs(adj - fd) <= RTOL * abs(fd)
    rec.close()
    ctx.close()


def test_torch_autograd_function():
    from aworkefficientgpualgorithmforlevelsetsegmentation_amd.adjoint import segment, soft_dice_loss
    vol, gt, phi0, (T, eps, alpha), K = _case(iters=30)
    ctx = Context(vol)
    dev = torch.device("cuda:0")
    p0 = torch.tensor(phi0, device=dev, requires_grad=True)
    theta = torch.tensor([T, eps, alpha], dtype=torch.float64, device=dev, requires_grad=True)
    mask = torch.tensor(gt, device=dev)
    phi = segment(ctx, p0, theta, dt=DT, iters=K)
    loss = soft_dice_loss(phi, mask)
    loss.backward()
    # the same through the record directly
    ref = Context(vol)
    ref.set_phi(phi0, 1.0)
    rec = ref.fwd_record(T, eps, alpha, DT, K)
    assert np.array_equal(phi.detach().cpu().numpy().view(np.uint32), ref.phi(0).view(np.uint32))
    _, dphi = _soft_dice_and_grad(ref.phi(0), gt)
    g0, th = rec.bwd(dphi.astype(np.float32))
    tg = theta.grad.cpu().numpy()
    _say(f"torch dtheta {tg} record {th}")
    assert np.allclose(tg, th, rtol=1e-4, atol=1e-7)
    pg = p0.grad.cpu().numpy()
    assert np.linalg.norm(pg - g0) <= 1e-4 * np.linalg.norm(g0)
    # a record outliving its context is detached, not dangling
    ref.close()
    with pytest.raises(Exception):
        rec.replay(phi0, T, eps, alpha, DT)
    rec.close()
    ctx.close()
    del phi, loss


def test_deferred_keep_step_records_the_placing_keep_steps_traversal(monkeypatch):
    """The update launch rebuilds the traversal that a deferred keep step
    leaves as kept directions in the order the placing keep step appends it,
    so both record V_k in the same order with the same decisions, and their
    replays agree bit for bit."""
    vol, _, phi0, (T, eps, alpha), K = _case(edge=(34, 30, 26), seed=91, iters=36)
    _say(f"records with the deferred and the placing keep step {vol.shape} K={K}")
    out = {}
    for defer in ("1", "0"):
        monkeypatch.setenv("LSSEG_DEFER_KEEP", defer)
        ctx = Context(vol)
        ctx.set_phi(phi0, 1.0)
        rec = ctx.fwd_record(T, eps, alpha, DT, K)
        n, lists, dec = rec.lists()
        out[defer] = (n, lists, dec, rec.replay(phi0, T, eps, alpha, DT), ctx.state())
        rec.close()
        ctx.close()
    a, b = out["1"], out["0"]
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    for x, y in zip(a[4], b[4]):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_record_past_the_stamp_wrap_matches_the_plain_run():
    """A recording run of 300 phantom iterations, whose list stays non-empty:
    the record's launches restart the emission stamp once in the call, with
    the last marking's compaction still pending.  The final state and the
    per-iteration counts equal a plain run's, and the replay of the record
    equals the forward."""
    w = synth.phantom(iterations=300)
    _say(f"record of the phantom, {w.iterations} iterations")
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    rec = ctx.fwd_record(w.T, w.eps, w.alpha, w.dt, w.iterations)
    n, _, _ = rec.lists()
    plain = Context(w.volume)
    plain.set_phi(w.phi0, w.band)
    plain.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    assert int(np.count_nonzero(n)) == w.iterations
    assert np.array_equal(np.asarray(n), plain.iteration_counts().n[:w.iterations])
    for a, b in zip(ctx.state(), plain.state()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    out = rec.replay(w.phi0, w.T, w.eps, w.alpha, w.dt)
    assert np.array_equal(out.view(np.uint32), ctx.phi(0).view(np.uint32)), "replay differs from the forward"
    rec.close()
    ctx.close()
    plain.close()

