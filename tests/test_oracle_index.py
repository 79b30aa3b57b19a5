"""Pins the oracle's index helpers (Coordinates.cu:65-96, 132-184; Cuda.cpp:188-200;
Segmenter.cpp:1363-1369) against closed forms and brute force."""
import itertools

import numpy as np

import oracle as O


def test_tau_is_a_bijection_matching_the_tile_formula():
    X, Y, Z = 48, 32, 16
    seen = np.zeros(X * Y * Z, bool)
    L = O.lib()
    for z, y, x in itertools.product(range(Z), range(Y), range(X)):
        t = L.lso_tau(X, Y, x, y, z)
        # closed form: tile-major, 16^3 tiles, x fastest inside a tile
        tile = ((z // 16) * (Y // 16) + y // 16) * (X // 16) + x // 16
        assert t == tile * 4096 + (z % 16) * 256 + (y % 16) * 16 + x % 16
        assert not seen[t]
        seen[t] = True
    assert seen.all()


def test_untau_inverts_tau_vectorised():
    X, Y, Z = 64, 48, 32
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    t = O.tau(X, Y, x, y, z)
    assert np.array_equal(np.sort(t.ravel()), np.arange(X * Y * Z))
    xx, yy, zz = O.untau(X, Y, t)
    assert np.array_equal(xx, x) and np.array_equal(yy, y) and np.array_equal(zz, z)


def test_tile_volume_roundtrip_and_agrees_with_tau():
    X, Y, Z = 32, 48, 16
    vol = np.arange(X * Y * Z, dtype=np.int64).reshape(Z, Y, X)
    flat = O.tile_volume(vol)
    z, y, x = 5, 37, 20
    assert flat[O.tau(X, Y, x, y, z)] == vol[z, y, x]
    assert np.array_equal(O.untile_volume(flat, X, Y, Z), vol)


def test_neighbor_clamps_to_self_at_the_boundary():
    X, Y, Z = 32, 32, 16
    L = O.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = int(rng.integers(X)), int(rng.integers(Y)), int(rng.integers(Z))
        t = L.lso_tau(X, Y, x, y, z)
        for e, (dx, dy, dz) in enumerate(O.DIRS):
            nx, ny, nz = x + dx, y + dy, z + dz
            want = t if not (0 <= nx < X and 0 <= ny < Y and 0 <= nz < Z) else L.lso_tau(X, Y, nx, ny, nz)
            assert L.lso_neighbor(X, Y, Z, t, e) == want


def test_align256_and_path_switch():
    L = O.lib()
    assert [L.lso_align256(v) for v in (0, 1, 255, 256, 257, 511, 512)] == [0, 256, 256, 256, 512, 512, 512]
    N = 16 ** 3  # 4096
    # few iff 7 * align256(n) < N
    assert O.path_is_few(0, N)
    assert O.path_is_few(256, N)        # 1792 < 4096
    assert O.path_is_few(512, N)        # 3584 < 4096
    assert not O.path_is_few(513, N)    # 7 * 768 = 5376
    assert O.path_is_few(4608, 32 ** 3)  # 7 * 4608 = 32256 < 32768
    assert not O.path_is_few(4609, 32 ** 3)
