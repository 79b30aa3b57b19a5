"""ctypes binding of the engine's C ABI (``include/lsseg/lsseg.h``).

``Context`` owns one ``lsseg_ctx``: a padded, tiled copy of the volume and all
device buffers.  Every method maps to one C entry point and raises
``LssegError`` on a non-OK status.  There is no fallback: if the HIP library
cannot be loaded, construction fails.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _build

OK = 0
ERR_INVALID_ARG, ERR_OUT_OF_MEMORY, ERR_DEVICE, ERR_CAPACITY, ERR_STATE, ERR_UNSUPPORTED = range(1, 7)
U8, I8, U16, I16, U32, I32, F32 = range(7)
MODE_REF, MODE_CLEAN = 0, 1
PHI_F32, PHI_I8 = 0, 1

DTYPE_CODE = {np.dtype(np.uint8): U8, np.dtype(np.int8): I8, np.dtype(np.uint16): U16,
              np.dtype(np.int16): I16, np.dtype(np.uint32): U32, np.dtype(np.int32): I32,
              np.dtype(np.float32): F32}

# every entry point declared in include/lsseg/lsseg.h
EXPORTS = (
    "lsseg_create", "lsseg_destroy", "lsseg_last_error", "lsseg_get_dims", "lsseg_seed_sphere",
    "lsseg_set_phi", "lsseg_iterate", "lsseg_iterate_reinit", "lsseg_get_iteration_counts",
    "lsseg_get_active_count", "lsseg_get_active", "lsseg_get_phi", "lsseg_get_mask",
    "lsseg_freeze", "lsseg_clear", "lsseg_get_state", "lsseg_set_state",
    "lsseg_get_device_buffers", "lsseg_checkpoint_save", "lsseg_checkpoint_restore",
    "lsseg_iterate_profiled",
    "lsseg_batch_create", "lsseg_batch_destroy", "lsseg_batch_last_error", "lsseg_batch_count",
    "lsseg_batch_volume", "lsseg_batch_iterate", "lsseg_batch_checkpoint_save", "lsseg_batch_checkpoint_restore",
    "lsseg_batch_get_tuning", "lsseg_batch_set_tuning", "lsseg_batch_autotune",
    "lsseg_fwd_record", "lsseg_record_destroy", "lsseg_record_info", "lsseg_record_get", "lsseg_fwd_replay",
    "lsseg_bwd", "lsseg_auto_parameters", "lsseg_segment_noalloc", "lsseg_probe_random_access",
    "lsseg_get_tuning", "lsseg_set_tuning", "lsseg_autotune",
)


class LssegError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"lsseg status {status}: {message}")
        self.status = status


class VolumeDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
                ("dtype", ctypes.c_int32), ("z_aniso", ctypes.c_float), ("phi_storage", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [("T", ctypes.c_float), ("eps", ctypes.c_float), ("alpha", ctypes.c_float),
                ("dt", ctypes.c_float), ("mode", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("sum_active", ctypes.c_uint64),
                ("sum_candidates", ctypes.c_uint64), ("sum_new_active", ctypes.c_uint64),
                ("active_final", ctypes.c_uint64), ("few_path_iters", ctypes.c_uint64),
                ("many_path_iters", ctypes.c_uint64), ("device_ms", ctypes.c_double),
                ("empty_skipped", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("update_blocks", ctypes.c_uint32), ("compact_blocks", ctypes.c_uint32),
                ("ms", ctypes.c_double), ("candidates", ctypes.c_uint32), ("keep_items", ctypes.c_uint32),
                ("compact_tile", ctypes.c_uint32), ("cached", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTimes(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("first_ms", ctypes.c_double),
                ("last_ms", ctypes.c_double), ("launches", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_live = None


def _track(obj):
    """Closes engine objects still alive at interpreter exit, before the HIP
    runtime shuts down."""
    global _live
    if _live is None:
        import atexit
        import weakref
        _live = weakref.WeakSet()

        def _close_all():
            for o in list(_live):
                try:
                    o.close()
                except Exception:
                    pass
        atexit.register(_close_all)
    _live.add(obj)


def library_path() -> str:
    return _build.hip_library()


def lib():
    """Load liblsseg_hip.so and declare its signatures."""
    global _lib
    if _lib is None:
        L = ctypes.CDLL(library_path())
        vp, u32, u64, i32, f32 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32,
                                  ctypes.c_float)
        P = ctypes.POINTER
        sig = {
            "lsseg_create": (i32, [P(VolumeDesc), vp, P(vp)]),
            "lsseg_destroy": (i32, [vp]),
            "lsseg_last_error": (ctypes.c_char_p, [vp]),
            "lsseg_get_dims": (i32, [vp, P(u32), P(u32), P(u32)]),
            "lsseg_seed_sphere": (i32, [vp, f32, f32, f32, u32]),
            "lsseg_set_phi": (i32, [vp, vp, f32]),
            "lsseg_iterate": (i32, [vp, P(Params), i32, P(Stats)]),
            "lsseg_iterate_reinit": (i32, [vp, P(Params)]),
            "lsseg_get_iteration_counts": (i32, [vp, vp, vp, vp, u64, P(u64)]),
            "lsseg_get_active_count": (i32, [vp, P(u64)]),
            "lsseg_get_active": (i32, [vp, vp, u64, P(u64)]),
            "lsseg_get_phi": (i32, [vp, i32, vp]),
            "lsseg_get_mask": (i32, [vp, vp]),
            "lsseg_freeze": (i32, [vp]),
            "lsseg_clear": (i32, [vp, i32]),
            "lsseg_get_state": (i32, [vp, vp, vp, vp, vp, u64, P(u64)]),
            "lsseg_set_state": (i32, [vp, vp, vp, vp, vp, u64]),
            "lsseg_get_device_buffers": (i32, [vp, P(vp), P(vp), P(vp)]),
            "lsseg_checkpoint_save": (i32, [vp, i32]),
            "lsseg_checkpoint_restore": (i32, [vp, i32]),
            "lsseg_iterate_profiled": (i32, [vp, P(Params), i32, P(Stats), P(KernelTimes), vp, u64]),
            "lsseg_batch_create": (i32, [P(VolumeDesc), P(vp), u32, P(vp)]),
            "lsseg_batch_destroy": (i32, [vp]),
            "lsseg_batch_last_error": (ctypes.c_char_p, [vp]),
            "lsseg_batch_count": (i32, [vp, P(u32)]),
            "lsseg_batch_volume": (i32, [vp, u32, P(vp)]),
            "lsseg_batch_iterate": (i32, [vp, P(Params), i32, P(Stats)]),
            "lsseg_batch_checkpoint_save": (i32, [vp, i32]),
            "lsseg_batch_checkpoint_restore": (i32, [vp, i32]),
            "lsseg_batch_get_tuning": (i32, [vp, P(Tuning)]),
            "lsseg_batch_set_tuning": (i32, [vp, P(Tuning)]),
            "lsseg_batch_autotune": (i32, [vp, P(Params), i32, P(Tuning)]),
            "lsseg_fwd_record": (i32, [vp, P(Params), i32, P(vp)]),
            "lsseg_record_destroy": (i32, [vp]),
            "lsseg_record_info": (i32, [vp, P(i32), P(u64)]),
            "lsseg_record_get": (i32, [vp, vp, vp, vp]),
            "lsseg_fwd_replay": (i32, [vp, vp, vp, P(Params), vp]),
            "lsseg_bwd": (i32, [vp, vp, vp, vp, vp]),
            "lsseg_auto_parameters": (i32, [vp, P(f32), P(f32), P(u64)]),
            "lsseg_segment_noalloc": (i32, [P(VolumeDesc), vp, vp, f32, P(Params), i32, vp, vp, vp, P(Stats)]),
            "lsseg_probe_random_access": (i32, [vp, u64, ctypes.c_uint32, i32, i32, vp, P(f32)]),
            "lsseg_get_tuning": (i32, [vp, P(Tuning)]),
            "lsseg_set_tuning": (i32, [vp, P(Tuning)]),
            "lsseg_autotune": (i32, [vp, P(Params), i32, P(Tuning)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


PROBE_LOAD, PROBE_LOAD7, PROBE_STORE = 0, 1, 2


def probe_random_access(buf_ptr: int, nbytes: int, count: int, mode: int, reps: int = 21, stream=None) -> float:
    """Median launch time (us) of `count` threads making one random 4-byte load
    (PROBE_LOAD), seven (PROBE_LOAD7) or one store (PROBE_STORE) each, in
    32-byte sectors of the device buffer at buf_ptr (nbytes bytes; a store
    overwrites words with zeros): the random-access floor the path's gathers
    run against."""
    us = ctypes.c_float(0.0)
    L = lib()
    st = L.lsseg_probe_random_access(ctypes.c_void_p(buf_ptr), nbytes, count, mode, reps,
                                     ctypes.c_void_p(stream) if stream else None, ctypes.byref(us))
    if st != 0:
        raise LssegError(st, L.lsseg_last_error(None).decode())
    return float(us.value)


def _p(a: np.ndarray):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class IterationCounts:
    n: np.ndarray   # active voxels entering each iteration
    c: np.ndarray   # candidate emissions before duplicate removal
    m: np.ndarray   # entries of the next active list


class Context:
    """One segmentation context (the C ABI's ``lsseg_ctx``)."""

    def __init__(self, volume_zyx: np.ndarray, z_aniso: float = 1.0, phi_storage: int = PHI_F32):
        vol = np.ascontiguousarray(volume_zyx)
        if vol.ndim != 3:
            raise ValueError("volume must be 3-D (z, y, x)")
        if vol.dtype not in DTYPE_CODE:
            raise ValueError(f"unsupported dtype {vol.dtype}")
        self.nz, self.ny, self.nx = (int(v) for v in vol.shape)
        self.dtype = vol.dtype
        self.phi_storage = phi_storage
        L = lib()
        desc = VolumeDesc(self.nx, self.ny, self.nz, DTYPE_CODE[vol.dtype], z_aniso, phi_storage)
        h = ctypes.c_void_p()
        st = L.lsseg_create(ctypes.byref(desc), _p(vol), ctypes.byref(h))
        if st != OK:
            raise LssegError(st, L.lsseg_last_error(None).decode())
        self._h = h
        self._owned = True
        self._dims()
        _track(self)

    @classmethod
    def _borrow(cls, handle, shape_zyx, dtype, owner, phi_storage=PHI_F32):
        """A view of a context owned by someone else (a batch volume)."""
        self = cls.__new__(cls)
        self.nz, self.ny, self.nx = (int(v) for v in shape_zyx)
        self.dtype = np.dtype(dtype)
        self.phi_storage = phi_storage
        self._h = handle
        self._owned = False
        self._owner = owner
        self._dims()
        return self

    def _dims(self):
        X, Y, Z = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self._check(lib().lsseg_get_dims(self._h, ctypes.byref(X), ctypes.byref(Y), ctypes.byref(Z)))
        self.X, self.Y, self.Z = X.value, Y.value, Z.value
        self.N = self.X * self.Y * self.Z

    # -- plumbing --
    def _check(self, st: int):
        if st != OK:
            raise LssegError(st, lib().lsseg_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                lib().lsseg_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- initialisation --
    def set_phi(self, phi0_zyx: np.ndarray, band: float = 1.0):
        phi = np.ascontiguousarray(phi0_zyx, dtype=np.float32)
        if phi.shape != (self.nz, self.ny, self.nx):
            raise ValueError("phi0 shape mismatch")
        self._check(lib().lsseg_set_phi(self._h, _p(phi), band))

    def set_phi_ptr(self, address: int, band: float = 1.0):
        """set_phi from raw memory (host or device) holding nz*ny*nx float32."""
        self._check(lib().lsseg_set_phi(self._h, ctypes.c_void_p(address), band))

    def phi_into_ptr(self, address: int, which: int = 0):
        """phi (unpadded float32) into raw memory (host or device)."""
        self._check(lib().lsseg_get_phi(self._h, which, ctypes.c_void_p(address)))

    def fwd_record(self, T, eps, alpha, dt, iters, mode=MODE_REF) -> "Record":
        prm = Params(T, eps, alpha, dt, mode)
        h = ctypes.c_void_p()
        self._check(lib().lsseg_fwd_record(self._h, ctypes.byref(prm), int(iters), ctypes.byref(h)))
        return Record(self, h)

    def seed_sphere(self, cx: float, cy: float, cz: float, radius: int):
        self._check(lib().lsseg_seed_sphere(self._h, cx, cy, cz, int(radius)))

    # -- iteration --
    def iterate(self, T, eps, alpha, dt, iters, mode=MODE_REF) -> Stats:
        prm = Params(T, eps, alpha, dt, mode)
        st = Stats()
        self._check(lib().lsseg_iterate(self._h, ctypes.byref(prm), int(iters), ctypes.byref(st)))
        return st

    def iterate_profiled(self, T, eps, alpha, dt, iters, mode=MODE_REF):
        """Returns (Stats, KernelTimes, per-launch milliseconds in launch order:
        update, marking, keep, compaction of iteration 0, then of iteration 1, ...)."""
        prm = Params(T, eps, alpha, dt, mode)
        st = Stats()
        kt = KernelTimes()
        per = np.zeros(4 * int(iters) + 1, np.float32)
        self._check(lib().lsseg_iterate_profiled(self._h, ctypes.byref(prm), int(iters), ctypes.byref(st),
                                                 ctypes.byref(kt), _p(per), per.size))
        return st, kt, per[:int(kt.launches)]

    def checkpoint_save(self, slot: int = 0):
        self._check(lib().lsseg_checkpoint_save(self._h, slot))

    def checkpoint_restore(self, slot: int = 0):
        self._check(lib().lsseg_checkpoint_restore(self._h, slot))

    def iterate_reinit(self, T, eps, alpha, dt, mode=MODE_REF):
        prm = Params(T, eps, alpha, dt, mode)
        self._check(lib().lsseg_iterate_reinit(self._h, ctypes.byref(prm)))

    def iteration_counts(self) -> IterationCounts:
        cnt = ctypes.c_uint64()
        self._check(lib().lsseg_get_iteration_counts(self._h, None, None, None, 0, ctypes.byref(cnt)))
        k = cnt.value
        n = np.zeros(max(k, 1), np.uint32)
        c = np.zeros(max(k, 1), np.uint64)
        m = np.zeros(max(k, 1), np.uint32)
        self._check(lib().lsseg_get_iteration_counts(self._h, _p(n), _p(c), _p(m), k, ctypes.byref(cnt)))
        return IterationCounts(n[:k], c[:k], m[:k])

    # -- readback --
    @property
    def active_count(self) -> int:
        n = ctypes.c_uint64()
        self._check(lib().lsseg_get_active_count(self._h, ctypes.byref(n)))
        return n.value

    def active(self) -> np.ndarray:
        n = self.active_count
        out = np.zeros(max(n, 1), np.uint32)
        got = ctypes.c_uint64()
        self._check(lib().lsseg_get_active(self._h, _p(out), len(out), ctypes.byref(got)))
        return out[:got.value].copy()

    def phi(self, which: int = 0) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx), np.float32)
        self._check(lib().lsseg_get_phi(self._h, which, _p(out)))
        return out

    def mask(self) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx), np.uint8)
        self._check(lib().lsseg_get_mask(self._h, _p(out)))
        return out

    def freeze(self):
        self._check(lib().lsseg_freeze(self._h))

    def auto_parameters(self):
        """(T, eps, count): mean and sample standard deviation of the source inside
        the current segmentation (Segmenter.cpp:357-478)."""
        T, eps, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_uint64()
        self._check(lib().lsseg_auto_parameters(self._h, ctypes.byref(T), ctypes.byref(eps), ctypes.byref(n)))
        return T.value, eps.value, n.value

    def clear(self, all_segmentations: bool = False):
        self._check(lib().lsseg_clear(self._h, 1 if all_segmentations else 0))

    def state(self):
        """Raw padded tiled state: (phi_read, phi_write, export, list)."""
        r = np.zeros(self.N, np.float32)
        w = np.zeros(self.N, np.float32)
        e = np.zeros(self.N, np.int8)
        n = self.active_count
        lst = np.zeros(max(n, 1), np.uint32)
        got = ctypes.c_uint64()
        self._check(lib().lsseg_get_state(self._h, _p(r), _p(w), _p(e), _p(lst), len(lst),
                                          ctypes.byref(got)))
        return r, w, e, lst[:got.value].copy()

    def set_state(self, phi_read, phi_write, export, lst):
        r = np.ascontiguousarray(phi_read, np.float32)
        w = np.ascontiguousarray(phi_write, np.float32)
        e = np.ascontiguousarray(export, np.int8)
        v = np.ascontiguousarray(lst, np.uint32)
        if r.size != self.N or w.size != self.N or e.size != self.N:
            raise ValueError("state arrays must have X*Y*Z entries")
        self._check(lib().lsseg_set_state(self._h, _p(r), _p(w), _p(e), _p(v) if v.size else None, v.size))

    def tuning(self) -> "Tuning":
        t = Tuning()
        self._check(lib().lsseg_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, update_blocks: int = 0, compact_blocks: int = 0, keep_items: int = 0, compact_tile: int = 0):
        """0 in a field: unchanged."""
        t = Tuning(update_blocks, compact_blocks, 0.0, 0, keep_items, compact_tile, 0)
        self._check(lib().lsseg_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, T: float, eps: float, alpha: float, dt: float, iters: int, mode=MODE_REF) -> "Tuning":
        """Times the launch shapes and widths on `iters` iterations from the
        current state, keeps the fastest and restores the state."""
        p = Params(T, eps, alpha, dt, mode)
        t = Tuning()
        self._check(lib().lsseg_autotune(self._h, ctypes.byref(p), iters, ctypes.byref(t)))
        return t

    def device_buffers(self):
        r, w, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._check(lib().lsseg_get_device_buffers(self._h, ctypes.byref(r), ctypes.byref(w), ctypes.byref(s)))
        return r.value, w.value, s.value


class Batch:
    """Many volumes advanced together (the C ABI's ``lsseg_batch``).  ``volume(i)``
    is a ``Context`` view of volume i for initialisation and readback."""

    def __init__(self, volumes_zyx, z_aniso: float = 1.0, phi_storage: int = PHI_F32):
        vols = [np.ascontiguousarray(v) for v in volumes_zyx]
        if not vols:
            raise ValueError("empty batch")
        for v in vols:
            if v.ndim != 3 or v.dtype not in DTYPE_CODE:
                raise ValueError("volumes must be 3-D arrays of a supported dtype")
        L = lib()
        descs = (VolumeDesc * len(vols))(*[VolumeDesc(v.shape[2], v.shape[1], v.shape[0], DTYPE_CODE[v.dtype],
                                                      z_aniso, phi_storage) for v in vols])
        ptrs = (ctypes.c_void_p * len(vols))(*[v.ctypes.data for v in vols])
        h = ctypes.c_void_p()
        st = L.lsseg_batch_create(descs, ptrs, len(vols), ctypes.byref(h))
        if st != OK:
            raise LssegError(st, L.lsseg_batch_last_error(None).decode())
        self._h = h
        self._vols = []
        _track(self)
        for i, v in enumerate(vols):
            c = ctypes.c_void_p()
            self._check(L.lsseg_batch_volume(self._h, i, ctypes.byref(c)))
            self._vols.append(Context._borrow(c, v.shape, v.dtype, self, phi_storage))

    def _check(self, st: int):
        if st != OK:
            raise LssegError(st, lib().lsseg_batch_last_error(self._h).decode())

    def __len__(self):
        return len(self._vols)

    def volume(self, i: int) -> Context:
        return self._vols[i]

    def close(self):
        if getattr(self, "_h", None):
            for v in self._vols:
                v._h = None
            lib().lsseg_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _params(params):
        return (Params * len(params))(*[Params(*p) if not isinstance(p, Params) else p for p in params])

    def iterate(self, params, iters):
        """params: one (T, eps, alpha, dt, mode) per volume.  Returns one Stats per volume."""
        if len(params) != len(self._vols):
            raise ValueError("one parameter set per volume")
        stats = (Stats * len(self._vols))()
        self._check(lib().lsseg_batch_iterate(self._h, self._params(params), int(iters), stats))
        return list(stats)

    def checkpoint_save(self, slot: int = 0):
        self._check(lib().lsseg_batch_checkpoint_save(self._h, slot))

    def checkpoint_restore(self, slot: int = 0):
        self._check(lib().lsseg_batch_checkpoint_restore(self._h, slot))

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(lib().lsseg_batch_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, update_blocks: int, compact_blocks: int, keep_items: int = 0, compact_tile: int = 0):
        """keep_items / compact_tile 0: unchanged."""
        t = Tuning(update_blocks, compact_blocks, 0.0, 0, keep_items, compact_tile, 0)
        self._check(lib().lsseg_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, params, iters) -> Tuning:
        t = Tuning()
        self._check(lib().lsseg_batch_autotune(self._h, self._params(params), int(iters), ctypes.byref(t)))
        return t


class Record:
    """A recorded run of the differentiable form (``lsseg_record``)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle
        _track(self)
        it, tot = ctypes.c_int32(), ctypes.c_uint64()
        ctx._check(lib().lsseg_record_info(self._h, ctypes.byref(it), ctypes.byref(tot)))
        self.iters, self.total_active = it.value, tot.value

    def close(self):
        if getattr(self, "_h", None):
            lib().lsseg_record_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lists(self):
        """(n per iteration, concatenated lists, decision bytes) on the host."""
        n = np.zeros(self.iters, np.uint32)
        lst = np.zeros(max(self.total_active, 1), np.uint32)
        dec = np.zeros(max(self.total_active, 1), np.uint8)
        self.ctx._check(lib().lsseg_record_get(self._h, _p(n), _p(lst), _p(dec)))
        return n, lst[:self.total_active], dec[:self.total_active]

    def replay_ptr(self, phi0_address: int, T, eps, alpha, dt, out_address: int, mode=MODE_REF):
        prm = Params(T, eps, alpha, dt, mode)
        self.ctx._check(lib().lsseg_fwd_replay(self.ctx._h, self._h, ctypes.c_void_p(phi0_address),
                                               ctypes.byref(prm), ctypes.c_void_p(out_address)))

    def replay(self, phi0_zyx: np.ndarray, T, eps, alpha, dt, mode=MODE_REF) -> np.ndarray:
        phi = np.ascontiguousarray(phi0_zyx, dtype=np.float32)
        out = np.zeros_like(phi)
        self.replay_ptr(phi.ctypes.data, T, eps, alpha, dt, out.ctypes.data, mode)
        return out

    def bwd_ptr(self, dphi_address: int, dphi0_address: int) -> np.ndarray:
        th = np.zeros(3, np.float64)
        self.ctx._check(lib().lsseg_bwd(self.ctx._h, self._h, ctypes.c_void_p(dphi_address),
                                        ctypes.c_void_p(dphi0_address), _p(th)))
        return th

    def bwd(self, dphi_zyx: np.ndarray):
        """(dL/dphi0, [dL/dT, dL/deps, dL/dalpha]) for dL/dphi of the final phi."""
        g = np.ascontiguousarray(dphi_zyx, dtype=np.float32)
        out = np.zeros_like(g)
        th = self.bwd_ptr(g.ctypes.data, out.ctypes.data)
        return out, th
