"""The oracle's fixed-point phi storage (the reference's LEVEL_SET_CHAR,
Config.hpp:72-80) on the CPU.

Pinned against the FP32 building blocks, each pinned elsewhere: one iteration
of the fixed-point state equals the FP32 update of the same read buffer
(lso_update) with every written value replaced by floor(127 v) / 127
(UpdateLevelSetVolume.cu:316-318 and the normalised read of Textures.cu:12);
the export byte and the temporal flag come from the unquantised value; the
marking runs on the stored values with those flags.  A run that quantised the
flag's operand instead would differ on the case built here.
"""
import numpy as np
import pytest

import oracle as O


def _stored(v):
    q = np.clip(np.floor(v.astype(np.float32) * np.float32(127.0)), -128, 127).astype(np.int32)
    out = q.astype(np.float32) / np.float32(127.0)
    out[q == -128] = -1.0
    return out


def _random_fixed_state(rng, N, n):
    k = rng.integers(-128, 128, N)
    R = (k.astype(np.float32) / np.float32(127.0)).astype(np.float32)
    R[k == -128] = -1.0
    W = R[rng.permutation(N)]
    E = rng.integers(-128, 128, N).astype(np.int8)
    V = rng.permutation(N)[:n].astype(np.uint32)
    return R, W, E, V


@pytest.mark.parametrize("n", [1500, 6000], ids=["few", "many"])
def test_one_iteration_is_the_fp32_update_quantised(n):
    rng = np.random.default_rng(21 + n)
    vol = rng.integers(0, 255, (32, 32, 32)).astype(np.uint8)
    st = O.State(vol, phi_storage=1)
    X, Y, Z = st.X, st.Y, st.Z
    R, W, E, V = _random_fixed_state(rng, st.N, n)
    st.set_state(R, W, E, V)
    T, eps, alpha, dt = 150.0, 60.0, 0.2, 0.5
    src = O.padded_tiled_source(vol)
    Wf, Ef, flags = O.update(X, Y, Z, src, R, T, eps, alpha, dt, V, W=W.copy(), E=E.copy())
    st.iterate(T, eps, alpha, dt, 1)
    r2, w2, e2 = st.state()
    expect = W.copy()
    expect[V] = _stored(Wf[V])
    assert np.array_equal(r2.view(np.uint32), expect.view(np.uint32)), "stored values"
    assert np.array_equal(w2.view(np.uint32), R.view(np.uint32)), "the other buffer is untouched"
    assert np.array_equal(e2[V], Ef[V]), "export byte of the unquantised value"
    assert np.array_equal(np.round(expect[V] * 127).astype(np.int8), e2[V])
    # the marking uses the update's flags (unquantised) on the stored values
    masks = O.mark(X, Y, Z, expect, flags, V)
    flags_q = (np.abs(expect[V] - R[V]) > np.float32(0.008)).astype(np.uint8)
    assert not np.array_equal(flags, flags_q), "the case must tell the two flags apart"
    lst_expect = (O.compact_few if O.path_is_few(n, st.N) else O.compact_many)(X, Y, Z, V, masks)
    assert np.array_equal(st.active(), lst_expect)


def test_values_stay_on_the_fixed_point_grid():
    rng = np.random.default_rng(4)
    vol = rng.integers(0, 255, (40, 36, 30)).astype(np.uint8)
    st = O.State(vol, phi_storage=1)
    phi0 = np.clip(np.linspace(-3, 3, vol.size).reshape(vol.shape), -4, 4).astype(np.float32)
    st.set_phi(phi0, 1.0)
    st.iterate(140.0, 50.0, 0.1, 0.5, 20)
    for buf in st.state()[:2]:
        k = buf * np.float32(127.0)
        assert np.all(np.abs(k - np.round(k)) < 1e-4)
