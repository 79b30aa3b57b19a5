"""A seeded sweep of random configurations against the oracle, through the C
ABI (needs a GPU).

Each case draws ragged dimensions (1 to 70 voxels per axis, so padding and
the clamp-to-self faces take every shape), one of the seven source types,
T / eps / alpha, a positive or negative time step, REF or CLEAN mode, an
initial band that is thin (few path) or wide (the dense path at first), and a
launch tuning of the context.  The bar is the parity bar: per-iteration
counts n_k, c_k, m_k equal, and after the run the list, both phi buffers (bit
for bit), the export bytes and the mask equal the oracle's.  The same draws
run in the int8 phi mode against the oracle's int8 mode, and in batches of
two to nine volumes of one source type, each volume with its own shape and
parameters, against the oracle volume by volume.
"""
import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import MODE_CLEAN, MODE_REF, PHI_I8, Batch, Context

pytestmark = pytest.mark.gpu

SOURCES = [
    (np.uint8, 0, 255), (np.int8, -128, 127), (np.uint16, 0, 4000), (np.int16, -2000, 2000),
    (np.uint32, 0, 100000), (np.int32, -100000, 100000), (np.float32, 0.0, 1.0),
]
TUNINGS = [(0, 0, 0, 0), (64, 16, 1, 1024), (2048, 64, 2, 256), (8, 24, 2, 1024), (4096, 2048, 1, 256)]
CASES = 48


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    nx, ny, nz = (int(v) for v in rng.integers(1, 71, 3))
    dtype, lo, hi = SOURCES[seed % len(SOURCES)]
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx].astype(np.float32)
    c = rng.uniform(0, 1, 3) * np.array([nx, ny, nz], np.float32)
    r = float(rng.uniform(2, 0.45 * max(nx, ny, nz) + 2))
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    inside = d < r
    span = float(hi) - float(lo)
    fg, bg = lo + 0.7 * span, lo + 0.3 * span
    vol = np.where(inside, fg, bg) + rng.normal(0, 0.05 * span, inside.shape)
    vol = np.clip(vol, lo, hi)
    vol = vol.astype(dtype) if np.dtype(dtype).kind == "f" else np.rint(vol).astype(dtype)
    r0 = float(rng.uniform(1, max(1.5, 0.3 * r)))
    phi0 = np.clip((d - r0) / 1.5, -4, 4).astype(np.float32)
    band = float(rng.choice([1.0, 1.0, 2.5, 4.5]))
    T = float(fg + rng.uniform(-0.1, 0.1) * span)
    eps = float(rng.uniform(0.1, 0.5) * span)
    alpha = float(rng.uniform(0.0, 0.3))
    dt = float(rng.choice([0.5, 0.5, 0.35, -0.3]))
    mode = int(rng.choice([MODE_REF, MODE_REF, MODE_CLEAN]))
    iters = int(rng.integers(20, 81))
    tuning = TUNINGS[int(rng.integers(0, len(TUNINGS)))]
    return vol, phi0, band, (T, eps, alpha, dt), mode, iters, tuning


def _same(ctx, ref, where):
    r, w, e, lst = ctx.state()
    rr, rw, re_ = ref.state()
    assert np.array_equal(lst, ref.active()), f"{where}: lists differ"
    assert np.array_equal(r.view(np.uint32), rr.view(np.uint32)), f"{where}: phi read differs"
    assert np.array_equal(w.view(np.uint32), rw.view(np.uint32)), f"{where}: phi write differs"
    assert np.array_equal(e, re_), f"{where}: export differs"
    assert np.array_equal(ctx.mask(), ref.mask()), f"{where}: masks differ"


def _say(seed, what, vol, band, prm, mode, iters, tuning):
    T, eps, alpha, dt = prm
    print(f"{what} {seed}: {vol.shape[::-1]} {vol.dtype} band {band} T={T:.4g} eps={eps:.4g} alpha={alpha:.3f} "
          f"dt={dt} mode={mode} iters={iters} tuning={tuning}", flush=True)


@pytest.mark.parametrize("phi", ["f32", "i8"])
@pytest.mark.parametrize("seed", range(CASES))
def test_random_configuration_matches_the_oracle(seed, phi):
    vol, phi0, band, (T, eps, alpha, dt), mode, iters, tuning = _case(seed)
    _say(seed, f"case ({phi})", vol, band, (T, eps, alpha, dt), mode, iters, tuning)
    ref = O.State(vol, phi_storage=1 if phi == "i8" else 0)
    ref.set_phi(phi0, band)
    n, c, m = ref.iterate(T, eps, alpha, dt, iters, mode=mode, threads=8)
    ctx = Context(vol, phi_storage=PHI_I8) if phi == "i8" else Context(vol)
    ctx.set_phi(phi0, band)
    if any(tuning):
        ctx.set_tuning(*tuning)
    ctx.iterate(T, eps, alpha, dt, iters, mode)
    cnt = ctx.iteration_counts()
    assert np.array_equal(cnt.n[:iters], n), f"case {seed}: n_k differ"
    assert np.array_equal(cnt.c[:iters], c), f"case {seed}: c_k differ"
    assert np.array_equal(cnt.m[:iters], m), f"case {seed}: m_k differ"
    _same(ctx, ref, f"case {seed} ({phi})")
    ctx.close()


@pytest.mark.parametrize("seed", range(12))
def test_random_batch_matches_the_oracle(seed):
    rng = np.random.default_rng(5000 + seed)
    count = int(rng.integers(2, 10))
    dtype = SOURCES[seed % len(SOURCES)][0]
    draws = []
    for i in range(5 * count):  # first draws of this source type, each with its own shape and parameters
        d = _case(7 * (100 + 11 * seed + i) + SOURCES.index(next(x for x in SOURCES if x[0] == dtype)))
        if d[0].dtype == np.dtype(dtype):
            draws.append(d)
        if len(draws) == count:
            break
    assert len(draws) == count
    iters = int(rng.integers(20, 61))
    print(f"batch {seed}: {count} volumes of {np.dtype(dtype).name}, {iters} iterations, shapes "
          f"{[d[0].shape[::-1] for d in draws]}", flush=True)
    batch = Batch([d[0] for d in draws])
    if seed % 3 == 1:
        batch.set_tuning(512, count + 8 + 8 * (seed % 4), 2, 1024)
    for i, d in enumerate(draws):
        batch.volume(i).set_phi(d[1], d[2])
    params = [(d[3][0], d[3][1], d[3][2], d[3][3], d[4]) for d in draws]
    batch.iterate(params, iters)
    for i, d in enumerate(draws):
        vol, phi0, band, (T, eps, alpha, dt), mode = d[:5]
        ref = O.State(vol)
        ref.set_phi(phi0, band)
        n, c, m = ref.iterate(T, eps, alpha, dt, iters, mode=mode, threads=8)
        cnt = batch.volume(i).iteration_counts()
        assert np.array_equal(cnt.n[:iters], n) and np.array_equal(cnt.c[:iters], c), f"batch {seed} volume {i}"
        assert np.array_equal(cnt.m[:iters], m), f"batch {seed} volume {i}"
        _same(batch.volume(i), ref, f"batch {seed} volume {i}")
    batch.close()
