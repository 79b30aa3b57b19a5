"""Synthetic workload generators: determinism, shapes and the statistics the
benchmark configurations specify."""
import numpy as np

from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth


def test_ellipsoids_workload():
    a = synth.ellipsoids()
    b = synth.ellipsoids()
    assert a.volume.shape == (128, 128, 128) and a.volume.dtype == np.uint8
    assert np.array_equal(a.volume, b.volume)
    gt = a.truth.astype(bool)
    inside, outside = a.volume[gt].astype(float), a.volume[~gt].astype(float)
    assert abs(inside.mean() - 180) < 0.5 and abs(outside.mean() - 60) < 0.5
    assert abs(outside.std() - 10) < 0.3
    assert (a.T, a.eps, a.alpha, a.dt, a.iterations) == (180.0, 30.0, 0.1, 0.5, 300)
    # phi0: signed distance to the radius-6 sphere, clamped to +-4
    assert a.phi0.min() == -4.0 and a.phi0.max() == 4.0
    c = [int(round(v)) for v in a.seed_center]
    assert a.phi0[c[2], c[1], int(a.seed_center[0] + 0.5)] == -4.0


def test_phantom_workload_small():
    w = synth.phantom(dims=(64, 72, 60))
    assert w.volume.shape == (60, 72, 64) and w.volume.dtype == np.float32
    x, y, z = (int(v) for v in w.seed_center)
    assert w.truth[z, y, x] == 1
    lab = w.truth.astype(bool)
    assert abs(np.median(w.volume[lab]) - 0.55) < 0.03


def test_vessels_workload_small():
    w = synth.vessels(edge=128, n_tubes=6, with_mask=True)
    assert w.volume.dtype == np.uint16
    m = w.truth.astype(bool)
    assert 0 < m.mean() < 0.2
    assert abs(w.volume[m].mean() - 1200) < 10 and abs(w.volume[~m].mean() - 300) < 5
    x, y, z = (int(v) for v in w.seed_center)
    assert m[z, y, x]


def test_batch_and_adjoint_workloads():
    b = synth.batch(count=3, iterations=10)
    for w in b:
        assert all(64 <= s <= 256 for s in w.volume.shape)
        assert 150 <= w.T <= 200 and 20 <= w.eps <= 40 and 0.05 <= w.alpha <= 0.2
    a = synth.adjoint(count=2, edge=32)
    assert a[0].volume.dtype == np.float32 and a[0].volume.max() <= 1.0
    assert a[0].truth.any()


def test_dice():
    a = np.zeros((4, 4), np.uint8)
    a[:2] = 1
    b = a.copy()
    assert synth.dice(a, b) == 1.0
    b[0] = 0
    assert abs(synth.dice(a, b) - 2 * 4 / 12) < 1e-12
