"""A reference-format project through the C++ lsseg::Segmenter, and the
auto-parameter statistics, against the oracle (needs a GPU).

The project (raw uint16 volume, *.par, *.gproj in the reference's bin/ layout)
is written by the test; segmenter_demo --project opens it the way the
reference application does (MainWindow::openProject, Engine::LoadVolume),
seeds, optionally runs CalculateSegmentationParameters, and requests updates.
The oracle is driven the same way; masks must be identical.
"""
import subprocess

import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import _build, formats, synth
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.engine import Context
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.segmenter import Parameters

pytestmark = pytest.mark.gpu


def _run_demo(gproj, seed, radius, requests, mask_out, auto):
    args = [_build.DEMO_BIN, "--project", gproj, *(str(v) for v in seed), str(radius), str(requests),
            str(mask_out), "1" if auto else "0"]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    kv = {}
    for line in res.stdout.splitlines():
        k, _, v = line.partition(" ")
        kv[k] = v
    iters = int(res.stdout.split("iterations")[1].split()[0])
    return kv, iters


@pytest.mark.parametrize("auto", [False, True], ids=["par", "auto"])
def test_project_through_cpp_segmenter(tmp_path, auto):
    print(f"project 64^3 uint16 vessels, auto parameters {auto}", flush=True)
    w = synth.vessels(edge=64, n_tubes=6)
    prm = Parameters(growLevel=1200.0, growWindow=250.0, timeStep=0.5, curvatureInfluence=0.05, numSubSteps=10)
    gproj = formats.write_project(str(tmp_path), "vessels64", w.volume, prm)
    out = tmp_path / "mask.raw"
    kv, iters = _run_demo(gproj, w.seed_center, 4, 6, out, auto)
    mask = np.fromfile(out, np.int8).reshape(w.volume.shape)
    ref = O.State(w.volume)
    ref.seed_sphere(*w.seed_center, 4)
    if auto:
        T_ref, eps_ref, n_ref = ref.auto_parameters()
        T, eps = float(kv["growLevel"]), float(kv["growWindow"])
        assert n_ref > 100
        assert abs(T - T_ref) <= 1e-4 * T_ref and abs(eps - eps_ref) <= 1e-4 * eps_ref
        # the parameter change restarts the active set on the first request
        ref.iterate_reinit(T, eps, 0.05, 0.5)
        ref.iterate(T, eps, 0.05, 0.5, iters - 1)
    else:
        assert float(kv["growLevel"]) == 1200.0 and int(kv["numSubSteps"]) == 10
        assert iters == 60
        ref.iterate(1200.0, 250.0, 0.05, 0.5, iters)
    assert np.array_equal(mask.astype(np.uint8), ref.mask())
    assert mask.sum() > 100


@pytest.mark.parametrize("kind", ["u8", "f32"])
def test_engine_auto_parameters(kind):
    if kind == "u8":
        w = synth.ellipsoids(edge=48)
        T, eps, alpha = 180.0, 30.0, 0.1
    else:
        w = synth.adjoint(count=1, edge=48)[0]
        T, eps, alpha = w.T, w.eps, w.alpha
    ctx = Context(w.volume)
    ctx.set_phi(w.phi0, w.band)
    ctx.iterate(T, eps, alpha, 0.5, 40)
    ref = O.State(w.volume)
    ref.set_phi(w.phi0, w.band)
    ref.iterate(T, eps, alpha, 0.5, 40)
    g = ctx.auto_parameters()
    r = ref.auto_parameters()
    print(f"{kind}: gpu {g} oracle {r}", flush=True)
    assert g[2] == r[2] and g[2] > 100
    assert abs(g[0] - r[0]) <= 1e-5 * abs(r[0]) and abs(g[1] - r[1]) <= 1e-4 * abs(r[1])
    ctx.close()
