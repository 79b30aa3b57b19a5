"""lsseg_segment_noalloc against the allocating entry points, through the C ABI
(needs a GPU).

The run that allocates nothing (its buffers in the library's static device
array) must give what lsseg_create + lsseg_set_phi + lsseg_iterate +
lsseg_get_phi + lsseg_get_mask give: phi bit for bit, the same mask and the same
statistics; on two calls in a row, the second with a smaller volume (the array
keeps the first call's contents), and the device's free memory is unchanged by
a call.  Oversized requests are refused before any device work.
"""
import ctypes

import numpy as np
import pytest

from aworkefficientgpualgorithmforlevelsetsegmentation_amd import engine, synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context, Params, Stats, VolumeDesc

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


class _Dev:
    """A device buffer through the engine library's HIP runtime (the test
    allocates; the entry point under test does not)."""

    def __init__(self, nbytes):
        L = engine.lib()
        self.ptr = ctypes.c_void_p()
        assert L.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 1))) == 0
        self.nbytes = nbytes

    def put(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert engine.lib().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert engine.lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        engine.lib().hipFree(self.ptr)


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert engine.lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _reference(w):
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    st = ctx.iterate(w.T, w.eps, w.alpha, w.dt, w.iterations)
    out = (ctx.phi(0), ctx.mask(), st)
    ctx.close()
    return out


def _noalloc(w, check_memory):
    L = engine.lib()
    nz, ny, nx = w.volume.shape
    src, phi0 = _Dev(w.volume.nbytes), _Dev(w.phi0.nbytes)
    phi_out, mask_out = _Dev(w.phi0.nbytes), _Dev(w.volume.size)
    src.put(w.volume)
    phi0.put(w.phi0.astype(np.float32))
    desc = VolumeDesc(nx, ny, nz, engine.DTYPE_CODE[w.volume.dtype], 1.0)
    prm = Params(w.T, w.eps, w.alpha, w.dt, engine.MODE_REF)
    st = Stats()
    assert L.hipDeviceSynchronize() == 0
    free0 = _free_bytes()
    rc = L.lsseg_segment_noalloc(ctypes.byref(desc), src.ptr, phi0.ptr, w.band, ctypes.byref(prm), w.iterations,
                                 phi_out.ptr, mask_out.ptr, None, ctypes.byref(st))
    assert rc == 0, L.lsseg_last_error(None)
    assert L.hipDeviceSynchronize() == 0
    if check_memory:
        assert _free_bytes() == free0, "the call changed the device's free memory"
    out = (phi_out.get((nz, ny, nx), np.float32), mask_out.get((nz, ny, nx), np.uint8), st)
    for b in (src, phi0, phi_out, mask_out):
        b.free()
    return out


def _same(a, b, where):
    phi_a, mask_a, st_a = a
    phi_b, mask_b, st_b = b
    assert np.array_equal(phi_a.view(np.uint32), phi_b.view(np.uint32)), f"{where}: phi differs"
    assert np.array_equal(mask_a, mask_b), f"{where}: mask differs"
    for f in ("iterations", "sum_active", "sum_candidates", "sum_new_active", "active_final", "few_path_iters",
              "many_path_iters"):
        assert getattr(st_a, f) == getattr(st_b, f), f"{where}: {f} {getattr(st_a, f)} vs {getattr(st_b, f)}"


def test_noalloc_run_equals_the_allocating_entry_points():
    big = synth.ellipsoids(edge=96, iterations=150)          # 96^3: 884,736 padded voxels
    small = synth.vessels(edge=48, n_tubes=3, iterations=90)  # smaller, another source type
    # more than 255 markings: the run clears its emission records and
    # restarts their stamp on the way
    long = synth.ellipsoids(edge=48, iterations=300, seed_offset=5)
    for i, w in enumerate((big, small, long)):
        _say(f"call {i}: {w.name} {w.volume.shape} {w.volume.dtype} {w.iterations} iterations")
        got = _noalloc(w, check_memory=i > 0)
        want = _reference(w)
        assert want[2].sum_active > 0 and want[1].any()
        _same(got, want, f"call {i}")


def test_noalloc_refuses_oversized_requests():
    L = engine.lib()
    prm = Params(100.0, 20.0, 0.1, 0.5, engine.MODE_REF)
    one = ctypes.c_void_p(1)  # never dereferenced: the checks come first
    desc = VolumeDesc(129, 128, 128, engine.U8, 1.0)           # pads to 144 x 128 x 128 > 2^21
    assert L.lsseg_segment_noalloc(ctypes.byref(desc), one, one, 1.0, ctypes.byref(prm), 10, one, one, None,
                                   None) == engine.ERR_UNSUPPORTED
    assert b"at most" in L.lsseg_last_error(None)
    desc = VolumeDesc(32, 32, 32, engine.U8, 1.0)
    assert L.lsseg_segment_noalloc(ctypes.byref(desc), one, one, 1.0, ctypes.byref(prm), 1 << 20, one, one, None,
                                   None) == engine.ERR_UNSUPPORTED
    assert L.lsseg_segment_noalloc(ctypes.byref(desc), None, one, 1.0, ctypes.byref(prm), 10, one, one, None,
                                   None) == engine.ERR_INVALID_ARG
