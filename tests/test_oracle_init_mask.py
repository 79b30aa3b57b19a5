"""Pins the oracle's seed initialisation (InitializeLevelSetVolume.cu:55-98 applied
twice, Segmenter.cpp:525-531; keep-flags InitializeActiveElementsVolume.cu:37-74),
band initialisation, mask predicate and export union
(InitializeLevelSetVolume.cu:151-169)."""
import numpy as np
import pytest

import oracle as O


def test_seed_sphere_closed_form():
    vol = np.zeros((32, 32, 32), np.uint8)
    s = O.State(vol)
    s.seed_sphere(16.7, 16.2, 16.9, 4)   # centre truncates to (16, 16, 16)
    R = s.phi(0)
    W = s.phi(1)
    # c(d) = -1 (d < 4), -1 + (d - 4) / 2 (d < 8), 1; read = clamp(2c - 1), write = c
    cases = {0: (-1.0, -1.0), 5: (-1.0, -0.5), 6: (-1.0, 0.0), 7: (0.0, 0.5), 8: (1.0, 1.0), 9: (1.0, 1.0)}
    for d, (r_want, w_want) in cases.items():
        assert R[16, 16, 16 + d] == pytest.approx(r_want, abs=1e-6), d
        assert W[16, 16, 16 + d] == pytest.approx(w_want, abs=1e-6), d
        assert R[16, 16 - d, 16] == pytest.approx(r_want, abs=1e-6), d


def test_seed_keep_flags_are_the_nonflat_voxels_in_tiled_order():
    vol = np.zeros((32, 32, 32), np.uint8)
    s = O.State(vol)
    s.seed_sphere(16, 16, 16, 4)
    R = s.phi(0)
    pad = np.pad(R, 1, mode="edge")
    core = pad[1:-1, 1:-1, 1:-1]
    flat = np.ones_like(R, bool)
    for ax in range(3):
        for sh in (-1, 1):
            nb = np.roll(pad, sh, axis=ax)[1:-1, 1:-1, 1:-1]
            flat &= np.abs(core - nb) <= np.float32(0.001)
    z, y, x = np.nonzero(~flat)
    want = np.sort(O.tau(32, 32, x, y, z))
    assert np.array_equal(s.active(), want)


def test_seed_anisotropy_stretches_z():
    vol = np.zeros((32, 32, 32), np.uint8)
    s = O.State(vol, z_aniso=2.0)
    s.seed_sphere(16, 16, 16, 4)
    R = s.phi(0)
    assert R[16, 16, 16 + 5] == -1.0   # 5 voxels in x: d = 5 -> c = -0.5 -> read -1
    assert R[16 + 5, 16, 16] == 1.0    # 5 voxels in z: d = 10 -> outside


def test_band_init_and_padding():
    rng = np.random.default_rng(0)
    vol = rng.integers(0, 255, (20, 18, 17)).astype(np.uint8)
    phi0 = rng.uniform(-3, 3, vol.shape).astype(np.float32)
    s = O.State(vol)
    assert (s.X, s.Y, s.Z) == (32, 32, 32)
    s.set_phi(phi0, 1.0)
    R, W, E = s.state()
    z, y, x = np.nonzero(np.abs(phi0) < 1.0)
    assert np.array_equal(s.active(), np.sort(O.tau(32, 32, x, y, z)))
    assert np.array_equal(s.phi(0), phi0) and np.array_equal(s.phi(1), phi0)
    full = O.untile_volume(R, 32, 32, 32)
    assert (full[20:] == 1.0).all() and (full[:, 18:] == 1.0).all() and (full[:, :, 17:] == 1.0).all()
    Eu = O.untile_volume(E, 32, 32, 32)[:20, :18, :17]
    assert np.array_equal(Eu, np.clip(np.floor(phi0 * np.float32(127)), -128, 127).astype(np.int8))


def test_signed_source_padding_bytes():
    # LoadVolumeHostPadded memsets -127 into every padding byte (Segmenter.cpp:1166),
    # so a padded int16 voxel reads 0x8181 = -32383; an update at such a voxel
    # has D = 1 exactly when T = -32383
    vol = np.full((10, 16, 16), -5, np.int16)
    s = O.State(vol)
    X, Y, Z = s.X, s.Y, s.Z
    x = np.arange(X, dtype=np.float32)
    ramp = np.broadcast_to(0.25 * (x - 8) + 0.125, (Z, Y, X)).astype(np.float32)
    R = O.tile_volume(np.ascontiguousarray(ramp))
    t = int(O.tau(X, Y, 8, 8, 12))   # z = 12 is padding
    out = {}
    for T in (-32383.0, -5.0):
        s.set_state(R, R, np.zeros(X * Y * Z, np.int8), np.array([t], np.uint32))
        s.iterate(T, 10.0, 0.0, 0.5, 1)
        out[T] = s.state()[0][t]
    # D = 1: 0.125 + 0.5 * -1 * 0.25 = 0.0 -> 0.01 after the rescale
    assert out[-32383.0] == pytest.approx(0.01, abs=1e-7)
    assert out[-5.0] == 1.0   # D = 1 - 32378 / 10 is hugely negative: grows to +1


def test_export_union_with_empty_frozen_counts_byte_one_as_inside():
    a = np.arange(-128, 128).astype(np.int8)
    r = O.add_export(a, np.full(256, 127, np.int8))
    inside = r <= 0
    # floor rounding in AddLevelSetKernel maps export byte 1 to 0 when the frozen
    # export is empty (127), so the saved mask is export <= 1
    assert np.array_equal(inside, a <= 1)
    # union with a fully inside frozen export (-127) is inside everywhere
    assert (O.add_export(a, np.full(256, -127, np.int8)) <= 0).all()


def test_mask_freeze_and_clear():
    vol = np.zeros((16, 16, 16), np.uint8)
    s = O.State(vol)
    phi = np.ones((16, 16, 16), np.float32)
    phi[4:8, 4:8, 4:8] = -1.0
    s.set_phi(phi, 0.5)
    m1 = s.mask()
    assert m1.sum() == 64
    s.freeze()
    s.clear(False)
    assert s.mask().sum() == 64       # frozen part survives a clear
    s.clear(True)
    assert s.mask().sum() == 0
