"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The oracle's C sources are built a third time with -fsanitize=address,undefined
(``make -C oracle sanitize``) and driven in a child process, which preloads the
ASan runtime and loads that build through LSORACLE_LIBRARY: set_phi, iterations
on the few and the dense path with the serial and the parallel compactions
(the sanitized build has no OpenMP, so both run serially), the seed, re-init,
masks, the stateless kernels, and the FP64 replay and adjoint.  Any invalid
access or undefined behaviour aborts the child with a report.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import numpy as np
import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth

w = synth.ellipsoids(edge=40, iterations=30)
st = O.State(w.volume)
st.set_phi(w.phi0, w.band)
st.iterate(w.T, w.eps, w.alpha, w.dt, 30, threads=2)
st.iterate(w.T, w.eps, w.alpha, w.dt, 5, mode=O.MODE_CLEAN)
assert st.n > 0 and st.mask().any()
st.freeze()
st.iterate_reinit(w.T - 5, w.eps, w.alpha, w.dt)
st.seed_sphere(12.0, 13.0, 14.0, 4)
st.iterate(w.T, w.eps, w.alpha, w.dt, 4)
st.clear(True)
n = 24
z, y, x = np.mgrid[0:n, 0:n, 0:n]
vol = ((x + y + z) % 7 * 30).astype(np.int16)
phi0 = np.clip((x - 11.5) / 3.0, -4, 4).astype(np.float32)
d = O.State(vol)
d.set_phi(phi0, 3.5)
nh, _, _ = d.iterate(100.0, 60.0, 0.2, 0.5, 12)
assert any(not O.path_is_few(int(k), d.N) for k in nh)
X, Y, Z = d.X, d.Y, d.Z
R = np.random.default_rng(1).uniform(-1, 1, d.N).astype(np.float32)
V = np.random.default_rng(2).permutation(d.N)[:3000].astype(np.uint32)
masks = O.mark(X, Y, Z, R, np.ones(3000, np.uint8), V)
assert np.array_equal(O.compact_few(X, Y, Z, V, masks), O.compact_few_par(X, Y, Z, V, masks, 1))
assert np.array_equal(O.compact_many(X, Y, Z, V, masks), O.compact_many_par(X, Y, Z, V, masks, 1))
a = synth.adjoint(count=1, edge=20, iterations=6)[0]
nk, lists, dec, _ = O.record_f32(a.volume, a.phi0, a.band, a.T, a.eps, a.alpha, a.dt, 6)
src = O.padded_tiled_source(a.volume)
Xa, Ya, Za = O.padded_dims(20, 20, 20)
out, hist = O.replay_f64(Xa, Ya, Za, src, a.T, a.eps, a.alpha, a.dt, nk, lists, dec, O.pad_tile_phi(a.phi0, 1.0),
                         keep_history=True)
g, th = O.adjoint_f64(Xa, Ya, Za, src, a.T, a.eps, a.alpha, a.dt, nk, lists, dec, hist, np.ones_like(out))
assert np.isfinite(th).all()
print("sanitized oracle ok")
'''


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_asan_and_ubsan():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["LSORACLE_LIBRARY"] = os.path.join(ROOT, "oracle", "_build", "liblsoracle_sanitize.so")
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    r = subprocess.run([sys.executable, "-c", CHILD], env=env, capture_output=True, text=True, timeout=600,
                       cwd=ROOT)
    assert r.returncode == 0, f"sanitized oracle failed ({r.returncode}):\n{r.stdout[-3000:]}\n{r.stderr[-6000:]}"
    assert "sanitized oracle ok" in r.stdout
