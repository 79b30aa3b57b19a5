"""The reference-shaped host APIs on the GPU: the C++ lsseg::Segmenter (through
its demo program) and the Python Segmenter mirror, against the oracle driven
the same way (seed sphere, numSubSteps iterations per request)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import Parameters, Segmenter, _build, synth

pytestmark = pytest.mark.gpu


def test_cpp_segmenter_matches_oracle(tmp_path):
    print("C++ Segmenter demo, 64^3 uint8, 20 requests x 10 sub-steps", flush=True)
    w = synth.ellipsoids(edge=64)
    raw = tmp_path / "vol.raw"
    w.volume.tofile(raw)
    out = tmp_path / "mask.raw"
    cx, cy, cz = w.seed_center
    args = [_build.DEMO_BIN, str(raw), "64", "64", "64", "1", "0", str(cx), str(cy), str(cz), "4",
            "180", "30", "0.5", "0.1", "10", "20", str(out)]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "iterations 200" in res.stdout
    mask = np.fromfile(out, np.int8).reshape(64, 64, 64)
    ref = O.State(w.volume)
    ref.seed_sphere(cx, cy, cz, 4)
    ref.iterate(180, 30, 0.1, 0.5, 200)
    assert np.array_equal(mask.astype(np.uint8), ref.mask())
    assert mask.sum() > 1000


def test_python_segmenter_mirror():
    print("Python Segmenter mirror with a parameter change", flush=True)
    w = synth.ellipsoids(edge=48)
    seg = Segmenter(Parameters(growLevel=180, growWindow=30, timeStep=0.5, curvatureInfluence=0.1,
                               numSubSteps=10))
    seg.LoadVolume(w.volume)
    seg.InitializeSeed(w.seed_center, 4)
    ref = O.State(w.volume)
    ref.seed_sphere(*w.seed_center, 4)
    assert seg.GetActiveElementCount() == ref.n
    for _ in range(3):
        seg.RequestUpdateSegmentation()
    ref.iterate(180, 30, 0.1, 0.5, 30)
    assert seg.GetActiveElementCount() == ref.n
    seg.parameters.growLevel = 170.0          # restart of the active set
    seg.RequestUpdateSegmentation()
    ref.iterate_reinit(170, 30, 0.1, 0.5)
    ref.iterate(170, 30, 0.1, 0.5, 9)
    assert seg.num_iterations == 10
    assert np.array_equal(seg.GetSaveableSegmentation().astype(np.uint8), ref.mask())
    assert seg.IsSegmentationInProgress() and not seg.IsSegmentationFinished()
    seg.RequestPauseSegmentation()
    assert not seg.IsSegmentationInProgress()
    seg.UnloadVolume()
