"""Check mode and the out-of-memory paths of the C ABI (needs a GPU).

- LSSEG_CHECKS=1 verifies every new list after its compaction (no entry twice,
  every entry where `pos` says, every traversal entry in the list): runs that
  take the few and the dense path raise no error and give the same states as
  unchecked runs.
- LSSEG_FAIL_ALLOC=k fails the k-th device allocation of a context creation:
  every k makes lsseg_create return LSSEG_ERR_OUT_OF_MEMORY with a message and
  no leaked device memory, and the library works normally afterwards.
- lsseg_probe_random_access (the measurement aid behind the bench line's
  random-access floor) times its three modes, writes only zeros, and refuses
  bad arguments.
"""
import ctypes

import numpy as np
import pytest

# torch ships its own HIP runtime and must provide the process's runtime: it is
# imported here, at collection, before any test loads the engine.
torch = pytest.importorskip("torch")

from aworkefficientgpualgorithmforlevelsetsegmentation_amd import engine, synth  # noqa: E402
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Batch, Context, LssegError  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _dense():
    n = 32
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    vol = ((x + y + z) % 7 * 30).astype(np.uint8)
    phi0 = np.clip((x - 15.5) / 3.0, -4, 4).astype(np.float32)
    return synth.Workload("dense", vol, phi0, 100.0, 60.0, 0.2, 0.5, 45, band=3.5)


@pytest.mark.parametrize("case", ["ellipsoids", "dense", "vessels"])
def test_check_mode_passes_and_changes_nothing(monkeypatch, case):
    w = {"ellipsoids": lambda: synth.ellipsoids(edge=64, iterations=100), "dense": _dense,
         "vessels": lambda: synth.vessels(edge=64, n_tubes=4, iterations=150)}[case]()
    out = {}
    for checks in ("0", "1"):
        monkeypatch.setenv("LSSEG_CHECKS", checks)
        _say(f"{case}: LSSEG_CHECKS={checks}")
        ctx = Context(w.volume)
        ctx.set_phi(w.phi0, w.band)
        st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
        st2 = ctx.iterate(w.T, w.eps, w.alpha, w.dt, 7)  # not a multiple of the graph chunk
        out[checks] = (st.sum_active, st2.sum_active, ctx.state())
        ctx.close()
    assert out["0"][:2] == out["1"][:2]
    for a, b in zip(out["0"][2], out["1"][2]):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert engine.lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_every_allocation_failure_of_create_is_reported_and_cleaned_up(monkeypatch):
    vol = np.zeros((160, 160, 160), np.uint16)  # a context of about 200 MB: a leak of any volume buffer shows
    L = engine.lib()
    assert L.hipDeviceSynchronize() == 0
    ok = Context(vol)
    ok.close()
    free0 = _free_bytes()
    failures = 0
    for k in range(1, 64):
        monkeypatch.setenv("LSSEG_FAIL_ALLOC", str(k))
        try:
            ctx = Context(vol)
        except LssegError as e:
            assert e.status == engine.ERR_OUT_OF_MEMORY, (k, str(e))
            assert "failed" in str(e)
            failures += 1
            continue
        ctx.close()  # k is past the last allocation: creation succeeds
        break
    monkeypatch.delenv("LSSEG_FAIL_ALLOC")
    _say(f"{failures} allocation points of lsseg_create failed cleanly")
    assert failures >= 20
    assert L.hipDeviceSynchronize() == 0
    leaked = free0 - _free_bytes()
    assert leaked < (32 << 20), f"{leaked} bytes of device memory not returned"
    # a batch whose second volume fails to allocate
    monkeypatch.setenv("LSSEG_FAIL_ALLOC", "3")
    with pytest.raises(LssegError) as e:
        Batch([np.zeros((32, 32, 32), np.uint8)] * 2)
    assert e.value.status == engine.ERR_OUT_OF_MEMORY
    monkeypatch.delenv("LSSEG_FAIL_ALLOC")
    w = synth.ellipsoids(edge=48, iterations=20)
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    assert ctx.iterate(w.T, w.eps, w.alpha, w.dt, 20).sum_active > 0
    ctx.close()


def test_random_access_probe():
    buf = torch.zeros(1 << 20, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    nbytes = buf.numel() * 4
    for mode in (engine.PROBE_LOAD, engine.PROBE_LOAD7, engine.PROBE_STORE):
        _say(f"probe mode {mode}: 10000 threads over {nbytes} bytes")
        us = engine.probe_random_access(buf.data_ptr(), nbytes, 10000, mode, reps=3)
        assert 0.0 < us < 1e5, us
    assert float(buf.abs().sum().item()) == 0.0  # the stores write zeros, the loads nothing
    for bad in ((nbytes, 0, 0), (100, 10, 0), (nbytes, 10, 3), (nbytes - 2, 10, 0)):
        with pytest.raises(LssegError):
            engine.probe_random_access(buf.data_ptr(), bad[0], bad[1], bad[2], reps=3)
