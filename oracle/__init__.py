"""ctypes binding of the CPU oracle (oracle/lsoracle.c).

TEST INFRASTRUCTURE ONLY: the test suite, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use this module, as the checker or the
timed CPU baseline.  The engine package never imports it.

All per-voxel arrays exchanged here are padded (multiples of 16) and stored in
the 16^3-tiled order of the reference (Coordinates.cu:65-96); ``tau`` /
``untau`` convert.  Lists hold tau values.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LSORACLE_LIBRARY selects another build of the same sources (the sanitizer
# test loads the ASan/UBSan build through it)
_LIB_PATH = os.environ.get("LSORACLE_LIBRARY") or os.path.join(_HERE, "_build", "liblsoracle.so")
# the same sources built with -march=native on the machine that runs them (the
# CPU baseline); the portable build above is what the tests load
_NATIVE_PATH = os.path.join(_HERE, "_build", "liblsoracle_native.so")

U8, I8, U16, I16, U32, I32, F32 = range(7)
MODE_REF, MODE_CLEAN = 0, 1
NZ, NY, NX, SELF, PX, PY, PZ = range(7)
DIRS = ((0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))

NP_DTYPE = {U8: np.uint8, I8: np.int8, U16: np.uint16, I16: np.int16,
            U32: np.uint32, I32: np.int32, F32: np.float32}
DTYPE_CODE = {np.dtype(v): k for k, v in NP_DTYPE.items()}

_libs = {}


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def build_native() -> str:
    """Compile the oracle for this machine's CPU (-march=native), for the CPU
    baseline; built where it runs, since the build machine's vector extensions
    need not exist on the machine that times it."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True)
    return _NATIVE_PATH


def lib(native: bool = False):
    """The oracle library (native: the -march=native build of build_native)."""
    L = _libs.get(native)
    if L is None:
        path = _NATIVE_PATH if native else _LIB_PATH
        if not os.path.exists(path):
            build_native() if native else build()
        L = ctypes.CDLL(path)
        c_i32, c_u32, c_i64, c_u64, c_f, c_d, vp = (ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64,
                                                    ctypes.c_uint64, ctypes.c_float, ctypes.c_double,
                                                    ctypes.c_void_p)
        L.lso_tau.restype = c_u32
        L.lso_tau.argtypes = [c_i32] * 5
        L.lso_neighbor.restype = c_u32
        L.lso_neighbor.argtypes = [c_i32, c_i32, c_i32, c_u32, c_i32]
        L.lso_align256.restype = c_u64
        L.lso_align256.argtypes = [c_u64]
        L.lso_path_is_few.restype = c_i32
        L.lso_path_is_few.argtypes = [c_u64, c_u64]
        L.lso_update_f32.restype = None
        L.lso_update_f32.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp, c_i64,
                                     c_f, c_f, c_f, c_f, c_i32]
        L.lso_update_f64.restype = None
        L.lso_update_f64.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp, c_i64,
                                     c_d, c_d, c_d, c_d, c_i32]
        L.lso_mark.restype = None
        L.lso_mark.argtypes = [c_i32, c_i32, c_i32, vp, vp, vp, c_i64, c_i32, vp, c_i32]
        L.lso_compact_few.restype = c_i64
        L.lso_compact_few.argtypes = [c_i32, c_i32, c_i32, vp, c_i64, vp, vp]
        L.lso_compact_many.restype = c_i64
        L.lso_compact_many.argtypes = [c_i32, c_i32, c_i32, vp, c_i64, vp, vp]
        L.lso_compact_few_par.restype = c_i64
        L.lso_compact_few_par.argtypes = [c_i32, c_i32, c_i32, vp, c_i64, vp, vp, vp, vp, c_i32]
        L.lso_compact_many_par.restype = c_i64
        L.lso_compact_many_par.argtypes = [c_i32, c_i32, c_i32, vp, c_i64, vp, vp, vp, c_i32]
        L.lso_create.restype = vp
        L.lso_create.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, c_f]
        L.lso_destroy.restype = None
        L.lso_destroy.argtypes = [vp]
        L.lso_dims.restype = None
        L.lso_dims.argtypes = [vp, vp, vp, vp]
        L.lso_set_phi.restype = c_i32
        L.lso_set_phi.argtypes = [vp, vp, c_f]
        L.lso_set_phi_storage.restype = None
        L.lso_set_phi_storage.argtypes = [vp, c_i32]
        L.lso_seed_sphere.restype = c_i32
        L.lso_seed_sphere.argtypes = [vp, c_f, c_f, c_f, c_u32]
        L.lso_iterate.restype = c_i32
        L.lso_iterate.argtypes = [vp, c_f, c_f, c_f, c_f, c_i32, c_i32, c_i32, vp, vp, vp]
        L.lso_active_count.restype = c_i64
        L.lso_active_count.argtypes = [vp]
        L.lso_get_active.restype = c_i64
        L.lso_get_active.argtypes = [vp, vp, c_i64]
        L.lso_get_state.restype = None
        L.lso_get_state.argtypes = [vp, vp, vp, vp]
        L.lso_set_state.restype = c_i32
        L.lso_set_state.argtypes = [vp, vp, vp, vp, vp, c_i64]
        L.lso_get_phi.restype = None
        L.lso_get_phi.argtypes = [vp, c_i32, vp]
        L.lso_get_mask.restype = None
        L.lso_get_mask.argtypes = [vp, vp]
        L.lso_reinit_active.restype = None
        L.lso_reinit_active.argtypes = [vp]
        L.lso_iterate_reinit.restype = None
        L.lso_iterate_reinit.argtypes = [vp, c_f, c_f, c_f, c_f]
        L.lso_freeze.restype = None
        L.lso_freeze.argtypes = [vp]
        L.lso_auto_parameters.restype = c_i64
        L.lso_auto_parameters.argtypes = [vp, vp, vp]
        L.lso_clear.restype = None
        L.lso_clear.argtypes = [vp, c_i32]
        L.lso_literal_begin.restype = c_i32
        L.lso_literal_begin.argtypes = [vp, c_i32]
        L.lso_iterate_literal.restype = c_i32
        L.lso_iterate_literal.argtypes = [vp, c_f, c_f, c_f, c_f, c_i32, c_i32, vp, vp, vp, vp]
        L.lso_add_export.restype = None
        L.lso_add_export.argtypes = [vp, vp, c_i64]
        _libs[native] = L
    return L


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def padded_dims(nx: int, ny: int, nz: int):
    return (nx + 15) // 16 * 16, (ny + 15) // 16 * 16, (nz + 15) // 16 * 16


def tau(X, Y, x, y, z):
    """16^3-tiled index; numpy-vectorised (Coordinates.cu:65-96)."""
    x, y, z = (np.asarray(a, dtype=np.int64) for a in (x, y, z))
    tile = ((z >> 4) * (Y >> 4) + (y >> 4)) * (X >> 4) + (x >> 4)
    return (tile * 4096 + ((z & 15) * 16 + (y & 15)) * 16 + (x & 15)).astype(np.int64)


def untau(X, Y, t):
    t = np.asarray(t, dtype=np.int64)
    TX, TY = X >> 4, Y >> 4
    tile, w = t >> 12, t & 4095
    tx, ty, tz = tile % TX, (tile // TX) % TY, tile // (TX * TY)
    return tx * 16 + (w & 15), ty * 16 + ((w >> 4) & 15), tz * 16 + (w >> 8)


def tile_volume(vol_zyx: np.ndarray) -> np.ndarray:
    """(Z, Y, X) array with dims multiples of 16 -> flat tiled array."""
    Z, Y, X = vol_zyx.shape
    v = vol_zyx.reshape(Z // 16, 16, Y // 16, 16, X // 16, 16)
    return np.ascontiguousarray(v.transpose(0, 2, 4, 1, 3, 5)).reshape(-1)


def untile_volume(flat: np.ndarray, X: int, Y: int, Z: int) -> np.ndarray:
    v = flat.reshape(Z // 16, Y // 16, X // 16, 16, 16, 16)
    return np.ascontiguousarray(v.transpose(0, 3, 1, 4, 2, 5)).reshape(Z, Y, X)


def update(X, Y, Z, src_tiled, R, T, eps, alpha, dt, V, W=None, E=None, threads=1):
    """Stateless update of every tau in V; returns (W, E, flags).  R's dtype picks
    the float32 or float64 instantiation."""
    L = lib()
    code = DTYPE_CODE[src_tiled.dtype]
    N = X * Y * Z
    W = R.copy() if W is None else W
    E = np.zeros(N, np.int8) if E is None else E
    V = np.ascontiguousarray(V, dtype=np.uint32)
    flags = np.zeros(max(len(V), 1), np.uint8)
    fn = L.lso_update_f64 if R.dtype == np.float64 else L.lso_update_f32
    fn(X, Y, Z, code, _p(src_tiled), _p(R), _p(W), _p(E), _p(flags), _p(V), len(V),
       T, eps, alpha, dt, threads)
    return W, E, flags[:len(V)]


def mark(X, Y, Z, R, flags, V, mode=MODE_REF, threads=1):
    L = lib()
    V = np.ascontiguousarray(V, dtype=np.uint32)
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    masks = np.zeros(max(len(V), 1), np.uint8)
    L.lso_mark(X, Y, Z, _p(np.ascontiguousarray(R, dtype=np.float32)), _p(flags), _p(V), len(V),
               mode, _p(masks), threads)
    return masks[:len(V)]


def compact_few(X, Y, Z, V, masks):
    L = lib()
    V = np.ascontiguousarray(V, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    out = np.zeros(max(7 * len(V), 1), np.uint32)
    m = L.lso_compact_few(X, Y, Z, _p(V), len(V), _p(masks), _p(out))
    return out[:m].copy()


def compact_many(X, Y, Z, V, masks):
    L = lib()
    V = np.ascontiguousarray(V, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    out = np.zeros(X * Y * Z, np.uint32)
    m = L.lso_compact_many(X, Y, Z, _p(V), len(V), _p(masks), _p(out))
    return out[:m].copy()


def compact_few_par(X, Y, Z, V, masks, threads=4):
    """compact_few with OpenMP (the CPU baseline's path); masks as the marking
    makes them (no out-of-bounds emission)."""
    L = lib()
    V = np.ascontiguousarray(V, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    out = np.zeros(max(7 * len(V), 1), np.uint32)
    tag = np.zeros(X * Y * Z, np.uint8)
    keep = np.zeros(max(len(V), 1), np.uint8)
    m = L.lso_compact_few_par(X, Y, Z, _p(V), len(V), _p(masks), _p(out), _p(tag), _p(keep), threads)
    assert not tag.any(), "the tag scratch must come back zero"
    return out[:m].copy()


def compact_many_par(X, Y, Z, V, masks, threads=4):
    L = lib()
    V = np.ascontiguousarray(V, dtype=np.uint32)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    out = np.zeros(X * Y * Z, np.uint32)
    flag = np.zeros(X * Y * Z, np.uint8)
    m = L.lso_compact_many_par(X, Y, Z, _p(V), len(V), _p(masks), _p(out), _p(flag), threads)
    assert not flag.any(), "the flag scratch must come back zero"
    return out[:m].copy()


def path_is_few(n, N):
    return bool(lib().lso_path_is_few(n, N))


class State:
    """Owning wrapper of lso_state (the oracle's Segmenter)."""

    def __init__(self, volume_zyx: np.ndarray, z_aniso: float = 1.0, phi_storage: int = 0, native: bool = False):
        """phi_storage 1: the reference's fixed-point (signed char) phi; native:
        run on the -march=native build (the CPU baseline)."""
        L = self._L = lib(native)
        vol = np.ascontiguousarray(volume_zyx)
        self.dtype_code = DTYPE_CODE[vol.dtype]
        self.nz, self.ny, self.nx = vol.shape
        self._keep = vol
        self._h = L.lso_create(self.nx, self.ny, self.nz, self.dtype_code, _p(vol), z_aniso)
        if not self._h:
            raise ValueError("lso_create failed")
        L.lso_set_phi_storage(self._h, 1 if phi_storage else 0)
        self.X, self.Y, self.Z = padded_dims(self.nx, self.ny, self.nz)
        self.N = self.X * self.Y * self.Z

    def close(self):
        if self._h:
            self._L.lso_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_phi(self, phi0_zyx: np.ndarray, band: float = 1.0):
        phi = np.ascontiguousarray(phi0_zyx, dtype=np.float32)
        assert phi.shape == (self.nz, self.ny, self.nx)
        self._L.lso_set_phi(self._h, _p(phi), band)

    def seed_sphere(self, cx, cy, cz, radius):
        self._L.lso_seed_sphere(self._h, cx, cy, cz, radius)

    def auto_parameters(self):
        """(T, eps, count): CalculateSegmentationParameters in the reference's float arithmetic."""
        T = ctypes.c_float()
        eps = ctypes.c_float()
        n = self._L.lso_auto_parameters(self._h, ctypes.byref(T), ctypes.byref(eps))
        return T.value, eps.value, int(n)

    def iterate(self, T, eps, alpha, dt, iters, mode=MODE_REF, threads=1):
        n = np.zeros(max(iters, 1), np.uint32)
        c = np.zeros(max(iters, 1), np.uint64)
        m = np.zeros(max(iters, 1), np.uint32)
        self._L.lso_iterate(self._h, T, eps, alpha, dt, mode, iters, threads, _p(n), _p(c), _p(m))
        return n[:iters], c[:iters], m[:iters]

    @property
    def n(self) -> int:
        return int(self._L.lso_active_count(self._h))

    def active(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.uint32)
        self._L.lso_get_active(self._h, _p(out), len(out))
        return out[:self.n].copy()

    def state(self):
        r = np.zeros(self.N, np.float32)
        w = np.zeros(self.N, np.float32)
        e = np.zeros(self.N, np.int8)
        self._L.lso_get_state(self._h, _p(r), _p(w), _p(e))
        return r, w, e

    def set_state(self, phi_read, phi_write, E, V):
        V = np.ascontiguousarray(V, dtype=np.uint32)
        rc = self._L.lso_set_state(self._h, _p(np.ascontiguousarray(phi_read, np.float32)),
                                 _p(np.ascontiguousarray(phi_write, np.float32)),
                                 _p(np.ascontiguousarray(E, np.int8)), _p(V), len(V))
        if rc != 0:
            raise ValueError("lso_set_state rejected the state")

    def phi(self, which: int = 0) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx), np.float32)
        self._L.lso_get_phi(self._h, which, _p(out))
        return out

    def mask(self) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx), np.uint8)
        self._L.lso_get_mask(self._h, _p(out))
        return out

    def literal_begin(self, initial_flags: bool = True):
        """Start the LITERAL replay (the shipped driver's scratch volumes) from
        the current state; see lso_literal_begin."""
        if self._L.lso_literal_begin(self._h, 1 if initial_flags else 0) != 0:
            raise MemoryError("lso_literal_begin failed")

    def iterate_literal(self, T, eps, alpha, dt, iters, threads=1):
        """(n, c, m, events[iters, 3]) of `iters` LITERAL iterations."""
        n = np.zeros(max(iters, 1), np.uint32)
        c = np.zeros(max(iters, 1), np.uint64)
        m = np.zeros(max(iters, 1), np.uint32)
        ev = np.zeros((max(iters, 1), 3), np.uint64)
        if self._L.lso_iterate_literal(self._h, T, eps, alpha, dt, iters, threads, _p(n), _p(c), _p(m), _p(ev)) != 0:
            raise RuntimeError("call literal_begin first")
        return n[:iters], c[:iters], m[:iters], ev[:iters]

    def iterate_reinit(self, T, eps, alpha, dt):
        self._L.lso_iterate_reinit(self._h, T, eps, alpha, dt)

    def freeze(self):
        self._L.lso_freeze(self._h)

    def clear(self, all_segmentations: bool = False):
        self._L.lso_clear(self._h, 1 if all_segmentations else 0)


def literal_diagnostic(volume_zyx, phi0_zyx, band, T, eps, alpha, dt, iters, threads=1, chunk=1,
                       initial_flags=True):
    """REF and LITERAL runs side by side from the same start (lso_set_phi):
    {"first_divergence": iteration whose next list differs, or None,
     "events": summed (stale valid flags compacted, candidates meeting a stale
     tag, repeated targets within a list), "first_event": first iteration with
     any event, or None, "iterations": iters, "many_iterations": dense-path
     iterations of the REF run}.  Lists are compared every `chunk` iterations
     (and counts every iteration)."""
    ref = State(volume_zyx)
    lit = State(volume_zyx)
    ref.set_phi(phi0_zyx, band)
    lit.set_phi(phi0_zyx, band)
    lit.literal_begin(initial_flags)
    out = {"first_divergence": None, "events": [0, 0, 0], "first_event": None, "iterations": iters,
           "many_iterations": 0}
    done = 0
    while done < iters:
        k = min(chunk, iters - done)
        n1, c1, m1 = ref.iterate(T, eps, alpha, dt, k, threads=threads)
        n2, c2, m2, ev = lit.iterate_literal(T, eps, alpha, dt, k, threads=threads)
        out["many_iterations"] += int(sum(1 for x in n1 if not path_is_few(int(x), ref.N)))
        hit = np.nonzero(ev.sum(axis=1))[0]
        if out["first_event"] is None and hit.size:
            out["first_event"] = done + int(hit[0])
        out["events"] = [a + int(b) for a, b in zip(out["events"], ev.sum(axis=0))]
        diff = np.nonzero((n1 != n2) | (c1 != c2) | (m1 != m2))[0]
        if diff.size:
            out["first_divergence"] = done + int(diff[0]) - (1 if diff[0] > 0 and n1[diff[0]] != n2[diff[0]] else 0)
            break
        if not np.array_equal(ref.active(), lit.active()):
            out["first_divergence"] = done + k - 1
            break
        done += k
    ref.close()
    lit.close()
    return out


def add_export(result: np.ndarray, add: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(result, dtype=np.int8).copy()
    a = np.ascontiguousarray(add, dtype=np.int8)
    lib().lso_add_export(_p(r), _p(a), r.size)
    return r


# ---- differentiable form (oracle/lsoracle_adjoint.c) -------------------------
def _adjoint_lib():
    L = lib()
    if not getattr(L, "_adjoint_declared", False):
        vp, c_i32, c_i64, c_f, c_d = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                      ctypes.c_double)
        L.lso_decisions_f32.restype = None
        L.lso_decisions_f32.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, vp, c_f, c_f, c_f, c_f, vp, c_i64, vp]
        L.lso_replay_f64.restype = None
        L.lso_replay_f64.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, c_d, c_d, c_d, c_d, c_i32, vp, vp, vp, vp,
                                     vp, vp]
        L.lso_adjoint_f64.restype = None
        L.lso_adjoint_f64.argtypes = [c_i32, c_i32, c_i32, c_i32, vp, c_d, c_d, c_d, c_d, c_i32, vp, vp, vp, vp,
                                      vp, vp, vp]
        L._adjoint_declared = True
    return L


def padded_tiled_source(vol_zyx: np.ndarray) -> np.ndarray:
    """The source padded to multiples of 16 (signed integer types pad with bytes
    0x81, everything else with zero bytes, Segmenter.cpp:1131-1204) and tiled."""
    vol = np.ascontiguousarray(vol_zyx)
    nz, ny, nx = vol.shape
    X, Y, Z = padded_dims(nx, ny, nz)
    signed_int = vol.dtype in (np.dtype(np.int8), np.dtype(np.int16), np.dtype(np.int32))
    raw = np.full((Z, Y, X * vol.itemsize), 0x81 if signed_int else 0, np.uint8)
    raw[:nz, :ny, :nx * vol.itemsize] = vol.view(np.uint8).reshape(nz, ny, nx * vol.itemsize)
    return tile_volume(raw.view(vol.dtype).reshape(Z, Y, X))


def pad_tile_phi(phi_zyx: np.ndarray, pad: float, dtype=np.float64) -> np.ndarray:
    nz, ny, nx = phi_zyx.shape
    X, Y, Z = padded_dims(nx, ny, nz)
    full = np.full((Z, Y, X), pad, dtype)
    full[:nz, :ny, :nx] = phi_zyx
    return tile_volume(full)


def unpad_untile(flat: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    X, Y, Z = padded_dims(nx, ny, nz)
    return untile_volume(flat, X, Y, Z)[:nz, :ny, :nx]


def decisions_f32(X, Y, Z, src_tiled, R, T, eps, alpha, dt, V):
    L = _adjoint_lib()
    V = np.ascontiguousarray(V, np.uint32)
    dec = np.zeros(max(len(V), 1), np.uint8)
    L.lso_decisions_f32(X, Y, Z, DTYPE_CODE[src_tiled.dtype], _p(src_tiled), _p(np.ascontiguousarray(R, np.float32)),
                        T, eps, alpha, dt, _p(V), len(V), _p(dec))
    return dec[:len(V)]


def record_f32(vol_zyx, phi0_zyx, band, T, eps, alpha, dt, iters, mode=MODE_REF):
    """The oracle's FP32 run with, per iteration, the list and the decisions of
    every update: (n per iteration, concatenated lists, decision bytes, final State)."""
    st = State(vol_zyx)
    st.set_phi(phi0_zyx, band)
    src = padded_tiled_source(vol_zyx)
    ns, lists, decs = [], [], []
    for _ in range(iters):
        V = st.active()
        R = st.state()[0]
        ns.append(len(V))
        lists.append(V)
        decs.append(decisions_f32(st.X, st.Y, st.Z, src, R, T, eps, alpha, dt, V))
        st.iterate(T, eps, alpha, dt, 1, mode=mode)
    cat = (lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t))
    return np.array(ns, np.uint32), cat(lists, np.uint32), cat(decs, np.uint8), st


def replay_f64(X, Y, Z, src_tiled, T, eps, alpha, dt, n, lists, dec, phi0_tiled, keep_history=False):
    L = _adjoint_lib()
    N = X * Y * Z
    n = np.ascontiguousarray(n, np.uint32)
    lists = np.ascontiguousarray(lists, np.uint32)
    dec = np.ascontiguousarray(dec, np.uint8)
    phi0 = np.ascontiguousarray(phi0_tiled, np.float64)
    out = np.zeros(N, np.float64)
    hist = np.zeros((len(n), N), np.float64) if keep_history else None
    L.lso_replay_f64(X, Y, Z, DTYPE_CODE[src_tiled.dtype], _p(src_tiled), T, eps, alpha, dt, len(n), _p(n),
                     _p(lists), _p(dec), _p(phi0), _p(out), _p(hist) if keep_history else None)
    return out, hist


def adjoint_f64(X, Y, Z, src_tiled, T, eps, alpha, dt, n, lists, dec, A_hist, dphiK_tiled):
    L = _adjoint_lib()
    N = X * Y * Z
    dphi0 = np.zeros(N, np.float64)
    th = np.zeros(3, np.float64)
    L.lso_adjoint_f64(X, Y, Z, DTYPE_CODE[src_tiled.dtype], _p(src_tiled), T, eps, alpha, dt, len(n),
                      _p(np.ascontiguousarray(n, np.uint32)), _p(np.ascontiguousarray(lists, np.uint32)),
                      _p(np.ascontiguousarray(dec, np.uint8)), _p(np.ascontiguousarray(A_hist, np.float64)),
                      _p(np.ascontiguousarray(dphiK_tiled, np.float64)), _p(dphi0), _p(th))
    return dphi0, th
