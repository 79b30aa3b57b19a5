"""Seeded synthetic workloads (volumes, initial phi, parameters).

The reference ships no input volumes for its published runs, so each workload
is generated from a seed by the C generators in ``csrc/synth.c`` (host code,
deterministic for a given seed and independent of the thread count).

Workloads (sizes and parameters of the five benchmark configurations):

* ``ellipsoids``  -- 128^3 uint8, two overlapping ellipsoids, T=180, eps=30,
  alpha=0.1, dt=0.5, 300 iterations;
* ``phantom``     -- 181x217x181 float32 four-class smooth phantom with a bias
  field, middle class, T=0.55, eps=0.12, alpha=0.15, dt=0.5, 1000 iterations;
* ``vessels``     -- 512^3 uint16 vessel tree of 40 tubes, T=1200, eps=250,
  alpha=0.05, dt=0.5, 2000 iterations;
* ``batch``       -- 64 uint8 ellipsoid volumes with uneven edges in [64, 256]
  and per-volume parameters, 400 iterations;
* ``adjoint``     -- 16 float32 96^3 ellipsoid volumes with ground-truth masks.

Every workload's initial phi is the signed distance to a seed sphere clamped
to +-4, and the initial active set is the band |phi| < 1.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _build

SEED_BASE = 0x5EED0000

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _build.synth_library()
        L = ctypes.CDLL(path)
        i32, u64, f32, vp = ctypes.c_int32, ctypes.c_uint64, ctypes.c_float, ctypes.c_void_p
        L.lsseg_synth_ellipsoids.restype = None
        L.lsseg_synth_ellipsoids.argtypes = [i32, i32, i32, u64, f32, f32, f32, f32, f32, i32, vp,
                                             vp, vp, f32, vp]
        L.lsseg_synth_sphere_sdf.restype = None
        L.lsseg_synth_sphere_sdf.argtypes = [i32, i32, i32, f32, f32, f32, f32, f32, vp]
        L.lsseg_synth_phantom.restype = i32
        L.lsseg_synth_phantom.argtypes = [i32, i32, i32, u64, f32, vp, vp, f32, f32, i32, f32,
                                          vp, vp, vp]
        L.lsseg_synth_vessels.restype = None
        L.lsseg_synth_vessels.argtypes = [i32, i32, i32, u64, i32, f32, i32, i32, f32, f32, f32,
                                          f32, f32, f32, f32, f32, vp, vp, vp]
        L.lsseg_synth_uniform.restype = ctypes.c_double
        L.lsseg_synth_uniform.argtypes = [u64, u64, u64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Workload:
    """One segmentation problem: source volume (z, y, x), initial phi, parameters."""
    name: str
    volume: np.ndarray
    phi0: np.ndarray
    T: float
    eps: float
    alpha: float
    dt: float
    iterations: int
    band: float = 1.0
    seed_center: tuple = (0.0, 0.0, 0.0)
    seed_radius: float = 0.0
    truth: np.ndarray | None = None
    extra: dict = field(default_factory=dict)

    @property
    def shape_xyz(self):
        z, y, x = self.volume.shape
        return (x, y, z)


def sphere_sdf(nx, ny, nz, center, radius, clampv=4.0) -> np.ndarray:
    phi = np.empty((nz, ny, nx), np.float32)
    lib().lsseg_synth_sphere_sdf(nx, ny, nz, float(center[0]), float(center[1]), float(center[2]),
                                 float(radius), float(clampv), _p(phi))
    return phi


def ellipsoid_volume(nx, ny, nz, seed, ellipsoids, bg=60.0, fg=180.0, sigma=10.0,
                     as_float_scale=None, truth=False):
    ell = np.ascontiguousarray(np.asarray(ellipsoids, np.float32).reshape(-1, 6))
    gt = np.empty((nz, ny, nx), np.uint8) if truth else None
    if as_float_scale is None:
        vol = np.empty((nz, ny, nx), np.uint8)
        lib().lsseg_synth_ellipsoids(nx, ny, nz, seed, bg, fg, sigma, 0.0, 255.0, len(ell), _p(ell),
                                     _p(vol), None, 1.0, _p(gt) if truth else None)
    else:
        vol = np.empty((nz, ny, nx), np.float32)
        lib().lsseg_synth_ellipsoids(nx, ny, nz, seed, bg, fg, sigma, 0.0, 255.0, len(ell), _p(ell),
                                     None, _p(vol), float(as_float_scale), _p(gt) if truth else None)
    return vol, gt


def _two_ellipsoids(nx, ny, nz, scale):
    """The two-ellipsoid layout of the 128^3 workload, scaled to other extents."""
    cx, cy, cz = nx / 2.0, ny / 2.0, nz / 2.0
    e1 = (cx - 7.5 * scale, cy, cz, 30 * scale, 22 * scale, 18 * scale)
    e2 = (cx + 7.5 * scale, cy, cz, 20 * scale, 28 * scale, 16 * scale)
    return [e1, e2]


def ellipsoids(edge: int = 128, iterations: int = 300, seed_offset: int = 0) -> Workload:
    """128^3 uint8: background 60, two ellipsoids of 180, noise sigma 10."""
    scale = edge / 128.0
    ells = _two_ellipsoids(edge, edge, edge, scale)
    vol, gt = ellipsoid_volume(edge, edge, edge, SEED_BASE + seed_offset, ells, truth=True)
    c = ells[0][:3]
    r = 6.0 * scale
    phi0 = sphere_sdf(edge, edge, edge, c, r)
    return Workload("ellipsoids", vol, phi0, 180.0, 30.0, 0.1, 0.5, iterations,
                    seed_center=c, seed_radius=r, truth=gt)


def phantom(iterations: int = 1000, dims=(181, 217, 181)) -> Workload:
    """181x217x181 float32 four-class phantom; segment the middle class."""
    nx, ny, nz = dims
    vol = np.empty((nz, ny, nx), np.float32)
    labels = np.empty((nz, ny, nx), np.uint8)
    q = np.array([0.30, 0.50, 0.85], np.float32)
    means = np.array([0.10, 0.35, 0.55, 0.80], np.float32)
    seed_xyz = np.zeros(3, np.float32)
    rc = lib().lsseg_synth_phantom(nx, ny, nz, SEED_BASE + 1, 4.0, _p(q), _p(means), 0.03, 0.10,
                                   2, 5.5, _p(vol), _p(labels), _p(seed_xyz))
    if rc != 0:
        raise RuntimeError("phantom generator found no seed location")
    c = tuple(float(v) for v in seed_xyz)
    phi0 = sphere_sdf(nx, ny, nz, c, 5.0)
    return Workload("phantom", vol, phi0, 0.55, 0.12, 0.15, 0.5, iterations,
                    seed_center=c, seed_radius=5.0, truth=(labels == 2).astype(np.uint8))


def vessels(iterations: int = 2000, edge: int = 512, n_tubes: int = 40, with_mask=False) -> Workload:
    """512^3 uint16 CT-like volume with a 40-tube vessel tree; seed at the root."""
    vol = np.empty((edge, edge, edge), np.uint16)
    mask = np.empty((edge, edge, edge), np.uint8) if with_mask else None
    root = np.zeros(3, np.float32)
    lib().lsseg_synth_vessels(edge, edge, edge, SEED_BASE + 2, n_tubes, 4.0, 60, 200, 2.0, 8.0, 6.0,
                              0.5, 300.0, 80.0, 1200.0, 50.0, _p(vol),
                              _p(mask) if with_mask else None, _p(root))
    c = tuple(float(v) for v in root)
    phi0 = sphere_sdf(edge, edge, edge, c, 4.0)
    return Workload("vessels", vol, phi0, 1200.0, 250.0, 0.05, 0.5, iterations,
                    seed_center=c, seed_radius=4.0, truth=mask)


def _u(stream, index):
    return float(lib().lsseg_synth_uniform(SEED_BASE + 3, stream, index))


def batch(count: int = 64, iterations: int = 400, edge_lo: int = 64, edge_hi: int = 256):
    """Uneven uint8 ellipsoid volumes with per-volume parameters."""
    out = []
    for i in range(count):
        u = [_u(i, k) for k in range(8)]
        dims = [int(round(math.exp(math.log(edge_lo) + (math.log(edge_hi) - math.log(edge_lo)) * u[k])))
                for k in range(3)]
        nx, ny, nz = dims
        scale = (0.5 + 0.7 * u[3]) * min(dims) / 128.0
        sigma = 5.0 + 10.0 * u[4]
        T = 150.0 + 50.0 * u[5]
        eps = 20.0 + 20.0 * u[6]
        alpha = 0.05 + 0.15 * u[7]
        ells = _two_ellipsoids(nx, ny, nz, scale)
        vol, gt = ellipsoid_volume(nx, ny, nz, SEED_BASE + 3 + 1000 * (i + 1), ells, sigma=sigma,
                                   truth=True)
        c = ells[0][:3]
        r = max(2.0, 6.0 * scale)
        phi0 = sphere_sdf(nx, ny, nz, c, r)
        out.append(Workload(f"batch{i}", vol, phi0, T, eps, alpha, 0.5, iterations,
                            seed_center=c, seed_radius=r, truth=gt))
    return out


def adjoint(count: int = 16, edge: int = 96, iterations: int = 200):
    """Float32 ellipsoid volumes rescaled to [0, 1] with ground-truth masks."""
    out = []
    scale = edge / 128.0
    for i in range(count):
        ells = _two_ellipsoids(edge, edge, edge, scale)
        vol, gt = ellipsoid_volume(edge, edge, edge, SEED_BASE + 4 + 1000 * (i + 1), ells,
                                   as_float_scale=1.0 / 255.0, truth=True)
        c = ells[0][:3]
        r = 6.0 * scale
        phi0 = sphere_sdf(edge, edge, edge, c, r)
        out.append(Workload(f"adjoint{i}", vol, phi0, 0.6, 0.15, 0.1, 0.5, iterations,
                            seed_center=c, seed_radius=r, truth=gt))
    return out


WORKLOADS = {"ellipsoids": ellipsoids, "phantom": phantom, "vessels": vessels}


def dice(a: np.ndarray, b: np.ndarray) -> float:
    a = a.astype(bool)
    b = b.astype(bool)
    s = a.sum() + b.sum()
    return 1.0 if s == 0 else float(2.0 * np.logical_and(a, b).sum() / s)


if __name__ == "__main__":  # pragma: no cover
    import time
    for name, fn in WORKLOADS.items():
        t = time.time()
        w = fn()
        print(name, w.volume.shape, w.volume.dtype, f"{time.time() - t:.2f}s", w.seed_center)
    os._exit(0)
