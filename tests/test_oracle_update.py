"""Pins the oracle's update rule (UpdateLevelSetVolume.cu:222-342) with values
computed by hand for chosen neighbourhoods, the sign and clamp branches, and
agreement of the float and double instantiations."""
import numpy as np
import pytest

import oracle as O

X = Y = Z = 16
C = (8, 8, 8)


def _field(u4, nz=None, ny=None, nx=None, px=None, py=None, pz=None, rest=None):
    """phi volume with u4 at C and the six given neighbour values."""
    vol = np.full((Z, Y, X), u4 if rest is None else rest, np.float32)
    x, y, z = C
    vol[z, y, x] = u4
    for val, (dx, dy, dz) in zip((nz, ny, nx, px, py, pz),
                                 ((0, 0, -1), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))):
        vol[z + dz, y + dy, x + dx] = u4 if val is None else val
    return O.tile_volume(vol)


def _src(value, dtype=np.uint8):
    return O.tile_volume(np.full((Z, Y, X), value, dtype))


def _one(R, src, T, eps, alpha, dt, at=C):
    t = int(O.tau(X, Y, *at))
    W, E, fl = O.update(X, Y, Z, src, R, T, eps, alpha, dt, np.array([t], np.uint32))
    return float(W[t]), int(E[t]), int(fl[0])


def test_ramp_at_target_intensity():
    # phi = 0.25 (x - 8) + 0.125 along x: |grad| = 0.25, Laplacian 0, no connectivity
    R = _field(0.125, nx=-0.125, px=0.375)
    # I = T: D = 1, dens = -1; speed = (1 - 0.2)(-1) = -0.8; delta = 0.5 * -0.8 * 0.25 = -0.1
    # 0.125 - 0.1 = 0.025 -> 1.01 * 0.025 + 0.01 = 0.03525
    v, e, f = _one(R, _src(180), 180, 30, 0.2, 0.5)
    assert v == pytest.approx(0.03525, abs=2e-7)
    assert e == 4  # floor(127 * 0.03525) = floor(4.47675)
    assert f == 1


def test_window_edge_freezes_the_data_term():
    # |I - T| = eps: D = 0, so only the rescale acts: 1.01 * 0.125 + 0.01
    R = _field(0.125, nx=-0.125, px=0.375)
    v, e, f = _one(R, _src(210), 180, 30, 0.2, 0.5)
    assert v == pytest.approx(0.13625, abs=2e-7)
    assert f == 1  # |0.13625 - 0.125| = 0.01125 > 0.008


def test_negative_time_step_flips_the_data_term():
    R = _field(0.125, nx=-0.125, px=0.375)
    # dens = +1: delta = +0.1 -> 0.225 -> 1.01 * 0.225 + 0.01 = 0.23725
    v, _, _ = _one(R, _src(180), 180, 30, 0.2, -0.5)
    assert v == pytest.approx(0.23725, abs=2e-7)


def test_connectivity_term_pulls_to_the_neighbour_mean():
    # all six neighbours 0.3, centre 0.1: central differences vanish (g = 0), so
    # delta = conn = 0.3 - 0.1 = 0.2 -> 0.3 -> 1.01 * 0.3 + 0.01 = 0.313
    R = _field(0.1, 0.3, 0.3, 0.3, 0.3, 0.3, 0.3)
    v, _, _ = _one(R, _src(0), 180, 30, 0.5, 0.5)
    assert v == pytest.approx(0.313, abs=2e-7)
    # with alpha = 0 the connectivity term is off: only the rescale acts
    v0, _, _ = _one(R, _src(0), 180, 30, 0.0, 0.5)
    assert v0 == pytest.approx(1.01 * 0.1 + 0.01, abs=2e-7)


def test_connectivity_needs_all_neighbours_within_tolerance():
    R = _field(0.1, 0.3, 0.3, 0.3, 0.3, 0.3, 0.41)  # |0.41 - 0.3| > 0.1
    # central z difference (0.41 - 0.3)/2 = 0.055; curvature = 3 * 0.4 + 0.11 = 1.31
    # I far from T: D = 1 - 180/30 = -5, dens = 5; speed = 0.5 * 5 + 0.5 * 1.31 = 3.155
    # delta = 0.5 * 3.155 * 0.055 = 0.0867625 -> 0.1867625 -> 1.01 x + 0.01 = 0.198630125
    v, _, _ = _one(R, _src(0), 180, 30, 0.5, 0.5)
    assert v == pytest.approx(0.198630125, abs=3e-7)


def test_clamps_saturate_at_minus_one():
    # u4 = -0.99, ramp of slope 0.5: delta = 0.5 * -0.8 * 0.5 = -0.2 -> clamp -1 -> 1.01 * -1 + 0.01 = -1
    R = _field(-0.99, nx=-1.49, px=-0.49)
    v, e, f = _one(R, _src(180), 180, 30, 0.2, 0.5)
    assert v == -1.0 and e == -127 and f == 1


def test_upper_clamp_and_rescale_fixed_point():
    R = _field(1.0)
    v, e, f = _one(R, _src(0), 180, 30, 0.2, 0.5)  # flat: g = 0, conn = 0 -> 1.01 + 0.01 -> 1
    assert v == 1.0 and e == 127 and f == 0


def test_boundary_neighbour_is_the_voxel_itself():
    # at x = 0 the -x neighbour is the voxel itself: dcx = (u5 - u4) / 2
    vol = np.full((Z, Y, X), 0.5, np.float32)
    vol[8, 8, 0] = 0.5
    vol[8, 8, 1] = 0.9
    R = O.tile_volume(vol)
    # dcx = 0.2, g = 0.2; curv = (0.9-0.5) - 0 = 0.4; D = 1 -> dens = -1
    # speed = 0.9 * -1 + 0.1 * 0.4 = -0.86; delta = 0.5 * -0.86 * 0.2 = -0.086
    # 0.414 -> 1.01 * 0.414 + 0.01 = 0.42814
    v, _, _ = _one(R, _src(180), 180, 30, 0.1, 0.5, at=(0, 8, 8))
    assert v == pytest.approx(0.42814, abs=2e-7)


@pytest.mark.parametrize("dtype,T,eps,val", [
    (np.int8, -20, 30, -35), (np.uint16, 1200, 250, 1300), (np.int16, -500, 100, -450),
    (np.uint32, 70000, 5000, 72000), (np.int32, -70000, 5000, -68000)])
def test_integer_sources_use_integer_distance(dtype, T, eps, val):
    # D = 1 - |I - T| / eps with integer arithmetic for |I - T|
    R = _field(0.125, nx=-0.125, px=0.375)
    D = 1.0 - abs(val - T) / eps
    want = 1.01 * (0.125 + 0.5 * (0.8 * -D) * 0.25) + 0.01
    v, _, _ = _one(R, _src(val, dtype), T, eps, 0.2, 0.5)
    assert v == pytest.approx(want, abs=3e-7)


def test_integer_sources_truncate_T_and_eps():
    R = _field(0.125, nx=-0.125, px=0.375)
    a = _one(R, _src(180), 180.9, 30.7, 0.2, 0.5)
    b = _one(R, _src(180), 180.0, 30.0, 0.2, 0.5)
    assert a == b


def test_float_source_uses_real_T_and_eps():
    R = _field(0.125, nx=-0.125, px=0.375)
    # D = 1 - |0.58 - 0.55| / 0.12 = 0.75; dens = -0.75; speed = 0.8 * -0.75 = -0.6
    # delta = 0.5 * -0.6 * 0.25 = -0.075 -> 0.05 -> 1.01 * 0.05 + 0.01 = 0.0605
    v, _, _ = _one(R, _src(0.58, np.float32), 0.55, 0.12, 0.2, 0.5)
    assert v == pytest.approx(0.0605, abs=5e-7)


def test_float_and_double_instantiations_agree():
    rng = np.random.default_rng(7)
    X2, Y2, Z2 = 32, 32, 16
    N = X2 * Y2 * Z2
    src = rng.integers(0, 256, N).astype(np.uint8)
    R = rng.uniform(-1, 1, N).astype(np.float32)
    V = rng.permutation(N)[:3000].astype(np.uint32)
    W32, _, f32 = O.update(X2, Y2, Z2, src, R, 150, 40, 0.3, 0.4, V)
    W64, _, f64 = O.update(X2, Y2, Z2, src, R.astype(np.float64), 150, 40, 0.3, 0.4, V)
    d = np.abs(W32[V].astype(np.float64) - W64[V])
    assert d.max() < 2e-6
    # only voxels whose |delta phi| sits on the threshold may disagree
    assert (f32 != f64).sum() <= 1
