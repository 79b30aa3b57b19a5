"""The five benchmark workloads at their own sizes, through the C ABI (needs a GPU).

Each configuration runs in full on the engine and on the oracle, from the same
seeded inputs:
  - ellipsoids 128^3 uint8, 300 iterations (in test_gpu_parity.py);
  - phantom 181x217x181 float32, all 1000 iterations;
  - vessels 512^3 uint16, all 2000 iterations (the list empties near
    iteration 1000; the engine stops launching then, the oracle runs on);
  - the batch of 64 uint8 volumes with edges log-uniform in [64, 256] and
    per-volume parameters, 400 iterations, every volume against its solo run
    and two volumes against the oracle;
  - the differentiable form on 16 float32 volumes of 96^3, K = 200: the record
    against the oracle's record, the GPU adjoint against the oracle's FP64
    adjoint (64^3 instance) and against central finite differences of the
    FP64 replay of the same recorded map at h = 1e-6.

Bar: per-iteration counts n_k, c_k, m_k equal; active lists, both phi buffers
(bit for bit) and export bytes equal at the checked iterations; final masks
identical (Dice 1.0 >= 0.9999).  Gradients within 1e-3 relative.
"""
import os
import time

import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Batch, Context

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
RTOL = 1e-3


def _say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _assert_same_state(ctx: Context, ref: O.State, where=""):
    r, w, e, lst = ctx.state()
    rr, rw, re_ = ref.state()
    rl = ref.active()
    assert np.array_equal(lst, rl), f"{where}: lists differ (gpu {len(lst)} vs oracle {len(rl)})"
    assert np.array_equal(r.view(np.uint32), rr.view(np.uint32)), f"{where}: phi read differs"
    assert np.array_equal(w.view(np.uint32), rw.view(np.uint32)), f"{where}: phi write differs"
    assert np.array_equal(e, re_), f"{where}: export differs"


def _full_run(w, legs):
    """Engine and oracle over the workload's iterations in `legs` (iteration
    counts per call); counts compared every leg, full states after each leg."""
    ctx = Context(w.volume)
    ref = O.State(w.volume)
    ctx.set_phi(w.phi0, w.band)
    ref.set_phi(w.phi0, w.band)
    _assert_same_state(ctx, ref, "init")
    assert sum(legs) == w.iterations
    done = 0
    total = 0
    for k in legs:
        t0 = time.perf_counter()
        st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, k)
        t1 = time.perf_counter()
        nh, ch, mh = ref.iterate(w.T, w.eps, w.alpha, w.dt, k, threads=THREADS)
        t2 = time.perf_counter()
        cnt = ctx.iteration_counts()
        assert np.array_equal(cnt.n, nh) and np.array_equal(cnt.c, ch) and np.array_equal(cnt.m, mh), \
            f"counts diverge in iterations {done}..{done + k}"
        assert st.sum_active == int(nh.sum())
        done += k
        total += int(nh.sum())
        _assert_same_state(ctx, ref, f"after {done} iterations")
        _say(f"{w.name}: {done} iterations equal (n={int(nh[-1])}, engine {t1 - t0:.2f} s incl. sync, "
             f"oracle {t2 - t1:.2f} s on {THREADS} threads)")
    m_gpu, m_ref = ctx.mask(), ref.mask()
    assert synth.dice(m_gpu, m_ref) >= 0.9999
    assert np.array_equal(m_gpu, m_ref)
    ctx.close()
    ref.close()
    return total


def test_phantom_all_1000_iterations():
    _say("phantom 181x217x181 float32, 1000 iterations")
    w = synth.phantom()
    total = _full_run(w, [250, 250, 250, 250])
    assert total > 10 ** 8


def test_vessels_512_all_2000_iterations():
    _say("vessels 512^3 uint16, 2000 iterations")
    w = synth.vessels()
    assert w.volume.shape == (512, 512, 512) and w.iterations == 2000
    total = _full_run(w, [200, 800, 1000])
    assert total > 10 ** 8


def _same_volume(a: Context, b: Context, where):
    ra, wa, ea, la = a.state()
    rb, wb, eb, lb = b.state()
    assert np.array_equal(la, lb), f"{where}: lists differ ({len(la)} vs {len(lb)})"
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32)), f"{where}: phi read differs"
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"{where}: phi write differs"
    assert np.array_equal(ea, eb), f"{where}: export differs"
    assert np.array_equal(a.mask(), b.mask()), f"{where}: mask differs"


def test_batch_64_volumes_400_iterations():
    ws = synth.batch()
    assert len(ws) == 64 and ws[0].iterations == 400
    _say(f"batch of {len(ws)} volumes ({sum(w.volume.size for w in ws) / 1e6:.0f} M voxels), 400 iterations")
    params = [(w.T, w.eps, w.alpha, w.dt, O.MODE_REF) for w in ws]
    batch = Batch([w.volume for w in ws])
    for i, w in enumerate(ws):
        batch.volume(i).set_phi(w.phi0, w.band)
    stats = batch.iterate(params, 400)
    _say("batch done; solo runs")
    for i, w in enumerate(ws):
        solo = Context(w.volume)
        solo.set_phi(w.phi0, w.band)
        st = solo.iterate(w.T, w.eps, w.alpha, w.dt, 400)
        assert (stats[i].sum_active, stats[i].sum_candidates, stats[i].active_final) == \
            (st.sum_active, st.sum_candidates, st.active_final), f"volume {i}"
        cb, cs = batch.volume(i).iteration_counts(), solo.iteration_counts()
        assert np.array_equal(cb.n, cs.n) and np.array_equal(cb.c, cs.c) and np.array_equal(cb.m, cs.m), i
        _same_volume(batch.volume(i), solo, f"volume {i}")
        solo.close()
    _say("every volume equals its solo run; two against the oracle")
    sizes = [w.volume.size for w in ws]
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        w = ws[i]
        ref = O.State(w.volume)
        ref.set_phi(w.phi0, w.band)
        nh, ch, mh = ref.iterate(w.T, w.eps, w.alpha, w.dt, 400, threads=THREADS)
        cb = batch.volume(i).iteration_counts()
        assert np.array_equal(cb.n, nh) and np.array_equal(cb.c, ch) and np.array_equal(cb.m, mh), i
        _assert_same_state(batch.volume(i), ref, f"volume {i} {w.volume.shape}")
        assert np.array_equal(batch.volume(i).mask(), ref.mask())
        ref.close()
    batch.close()


def _soft_dice_and_grad(phi, mask, temp=0.5):
    phi = phi.astype(np.float64)
    m = mask.astype(np.float64)
    s = 1.0 / (1.0 + np.exp(phi / temp))
    a = (s * m).sum()
    b = s.sum() + m.sum()
    return 1.0 - 2.0 * a / b, (-2.0 * (m * b - a) / (b * b)) * (-s * (1.0 - s) / temp)


def _same_record(n, lists, dec, on, olists, odec):
    """V_k and its decisions per iteration, both sorted by voxel."""
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


class _Fp64Map:
    """The FP64 replay of a recorded map and its soft-Dice loss (oracle)."""

    def __init__(self, w, n, lists, dec):
        nz, ny, nx = w.volume.shape
        self.w, self.n, self.lists, self.dec = w, n, lists, dec
        self.shape = (nx, ny, nz)
        self.X, self.Y, self.Z = O.padded_dims(nx, ny, nz)
        self.src = O.padded_tiled_source(w.volume)

    def loss(self, phi0, T, eps, alpha):
        out, _ = O.replay_f64(self.X, self.Y, self.Z, self.src, T, eps, alpha, self.w.dt, self.n, self.lists,
                              self.dec, O.pad_tile_phi(phi0, 1.0))
        return _soft_dice_and_grad(O.unpad_untile(out, *self.shape), self.w.truth)[0]


def test_adjoint_16x96_k200_record_fd_and_fp64():
    ws = synth.adjoint()
    assert len(ws) == 16 and ws[0].volume.shape == (96, 96, 96) and ws[0].iterations == 200
    K = 200
    h = 1e-6
    worst_theta = worst_dir = 0.0
    for i, w in enumerate(ws):
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        rec = ctx.fwd_record(w.T, w.eps, w.alpha, w.dt, K)
        n, lists, dec = rec.lists()
        if i < 2:
            on, olists, odec, ost = O.record_f32(w.volume, w.phi0, w.band, w.T, w.eps, w.alpha, w.dt, K)
            assert _same_record(n, lists, dec, on, olists, odec), f"volume {i}: record differs from the oracle's"
            assert np.array_equal(ctx.phi(0).view(np.uint32), ost.phi(0).view(np.uint32))
            ost.close()
        loss, dphi = _soft_dice_and_grad(ctx.phi(0), w.truth)
        g0, th = rec.bwd(dphi.astype(np.float32))
        fm = _Fp64Map(w, n, lists, dec)
        base = [w.T, w.eps, w.alpha]
        phi0 = w.phi0.astype(np.float64)
        for k in range(3):
            up, dn = list(base), list(base)
            up[k] += h
            dn[k] -= h
            fd = (fm.loss(phi0, *up) - fm.loss(phi0, *dn)) / (2 * h)
            err = abs(th[k] - fd) / abs(fd)
            worst_theta = max(worst_theta, err)
            assert err <= RTOL, (i, k, th[k], fd)
        d = g0.astype(np.float64) / np.linalg.norm(g0)
        fd = (fm.loss(phi0 + h * d, *base) - fm.loss(phi0 - h * d, *base)) / (2 * h)
        adj = float((g0.astype(np.float64) * d).sum())
        err = abs(adj - fd) / abs(fd)
        worst_dir = max(worst_dir, err)
        assert err <= RTOL, (i, adj, fd)
        _say(f"volume {i}: {int(n.sum())} recorded updates; theta {th}; FD(FP64, h=1e-6) worst so far "
             f"theta {worst_theta:.2e}, phi0 direction {worst_dir:.2e}")
        rec.close()
        ctx.close()
    # the FP64 host adjoint of the same recorded map on one 64^3 instance
    w = synth.adjoint(count=1, edge=64, iterations=K)[0]
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    rec = ctx.fwd_record(w.T, w.eps, w.alpha, w.dt, K)
    _, dphi = _soft_dice_and_grad(ctx.phi(0), w.truth)
    g0, th = rec.bwd(dphi.astype(np.float32))
    n, lists, dec = rec.lists()
    fm = _Fp64Map(w, n, lists, dec)
    out64, hist = O.replay_f64(fm.X, fm.Y, fm.Z, fm.src, w.T, w.eps, w.alpha, w.dt, n, lists, dec,
                               O.pad_tile_phi(w.phi0, 1.0), keep_history=True)
    _, dphi64 = _soft_dice_and_grad(O.unpad_untile(out64, *fm.shape), w.truth)
    g64, th64 = O.adjoint_f64(fm.X, fm.Y, fm.Z, fm.src, w.T, w.eps, w.alpha, w.dt, n, lists, dec, hist,
                              O.pad_tile_phi(dphi64, 0.0))
    del hist
    g64 = O.unpad_untile(g64, *fm.shape)
    for k in range(3):
        assert abs(th[k] - th64[k]) <= RTOL * abs(th64[k]), (k, th[k], th64[k])
    assert np.linalg.norm(g0 - g64) <= RTOL * np.linalg.norm(g64)
    _say(f"64^3 FP64 adjoint: theta gpu {th} fp64 {th64}")
    rec.close()
    ctx.close()
