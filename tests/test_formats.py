"""The reference's file formats on the CPU: raw volumes, *.par parameter files,
*.gproj projects and the per-type defaults.

Pinned three ways: a written project round-trips through the Python loaders;
the C++ loaders (segmenter_demo --parse, no device work) read the same values;
and, where the reference checkout is present, its own Bucky project and
default parameter files parse to the values they hold (the default table in
the loaders is checked against bin/defaults/*.par).  The oracle's
CalculateSegmentationParameters restatement is checked against numpy.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import _build, formats
from aworkefficientgpualgorithmforlevelsetsegmentation_amd.segmenter import Parameters

REF_BIN = "/root/reference/code/gpuseg/bin"


def _demo_parse(gproj):
    out = subprocess.run([_build.DEMO_BIN, "--parse", gproj], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    kv = {}
    for line in out.stdout.splitlines():
        k, _, v = line.partition(" ")
        kv[k] = v
    return kv


def _check_demo_matches(gproj):
    pr = formats.load_project(gproj)
    vol = formats.load_volume_raw(pr.volume)
    p = formats.project_parameters(pr, vol)
    kv = _demo_parse(gproj)
    assert os.path.realpath(kv["dataFile"]) == os.path.realpath(pr.volume.fileName)
    assert kv["numVoxels"] == f"{pr.volume.numVoxelsX} {pr.volume.numVoxelsY} {pr.volume.numVoxelsZ}"
    assert int(kv["numBytesPerVoxel"]) == pr.volume.numBytesPerVoxel
    assert int(kv["isSigned"]) == int(pr.volume.isSigned)
    # %.9g round-trips a float32
    assert np.float32(float(kv["minValue"])) == np.float32(vol.minValue)
    assert np.float32(float(kv["maxValue"])) == np.float32(vol.maxValue)
    for name in ("growLevel", "growWindow", "timeStep", "curvatureInfluence", "convergenceThreshold"):
        assert np.float32(float(kv[name])) == np.float32(getattr(p, name)), name
    assert int(kv["numSubSteps"]) == p.numSubSteps
    return pr, vol, p


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32])
def test_written_project_round_trips(tmp_path, dtype):
    rng = np.random.default_rng(5)
    vol = rng.integers(np.iinfo(dtype).min // 2, np.iinfo(dtype).max // 2, (9, 13, 17)).astype(dtype)
    prm = Parameters(growLevel=123.5, growWindow=45.25, timeStep=0.253333, curvatureInfluence=0.16,
                     numSubSteps=15, convergenceThreshold=3.0)
    gproj = formats.write_project(str(tmp_path), "case", vol, prm, z_anisotropy=1.5)
    pr, v, p = _check_demo_matches(gproj)
    assert np.array_equal(v.data, vol)
    assert v.minValue == float(np.float32(vol.min())) and v.maxValue == float(np.float32(vol.max()))
    assert pr.volume.zAnisotropy == 1.5
    assert (p.growLevel, p.growWindow, p.numSubSteps, p.convergenceThreshold) == (123.5, 45.25, 15, 3.0)
    assert np.float32(p.timeStep) == np.float32(0.253333)


def test_defaults_by_type():
    assert formats.default_parameters(255.0).growLevel == 150.0
    assert formats.default_parameters(256.0).growLevel == 790.0
    assert formats.default_parameters(4095.0).growWindow == 350.0
    assert formats.default_parameters(4096.0).growLevel == 15000.0
    with pytest.raises(ValueError):
        formats.default_parameters(65536.0)


def test_short_raw_file_is_an_error(tmp_path):
    (tmp_path / "v.raw").write_bytes(b"\x00" * 10)
    with pytest.raises(ValueError):
        formats.load_volume_raw(formats.VolumeFileDesc(str(tmp_path / "v.raw"), 4, 4, 4))


@pytest.mark.skipif(not os.path.isdir(REF_BIN), reason="reference checkout not present")
def test_reference_bucky_project():
    pr, vol, p = _check_demo_matches(os.path.join(REF_BIN, "projects", "Bucky.gproj"))
    assert (pr.volume.numVoxelsX, pr.volume.numVoxelsY, pr.volume.numVoxelsZ) == (32, 32, 32)
    assert pr.volume.numBytesPerVoxel == 1 and not pr.volume.isSigned
    assert vol.data.dtype == np.uint8 and vol.data.shape == (32, 32, 32)
    assert (p.growLevel, p.growWindow, p.numSubSteps) == (140.0, 80.0, 15)
    assert np.float32(p.timeStep) == np.float32(0.253333) and np.float32(p.curvatureInfluence) == np.float32(0.16)


@pytest.mark.skipif(not os.path.isdir(REF_BIN), reason="reference checkout not present")
@pytest.mark.parametrize("name,max_value", [("UnsignedChar", 255.0), ("UnsignedShort", 4095.0),
                                            ("UnsignedShortHDR", 65535.0)])
def test_default_table_matches_reference_defaults(name, max_value):
    ref = formats.load_parameters(os.path.join(REF_BIN, "defaults", f"{name}.par"))
    table = formats.default_parameters(max_value)
    for k in ("growLevel", "growWindow", "timeStep", "curvatureInfluence", "numSubSteps", "convergenceThreshold"):
        assert np.float32(getattr(ref, k)) == np.float32(getattr(table, k)), (name, k)


def test_oracle_auto_parameters_against_numpy():
    rng = np.random.default_rng(11)
    vol = rng.integers(0, 4000, (21, 19, 23)).astype(np.uint16)
    st = O.State(vol)
    st.seed_sphere(11.0, 9.0, 10.0, 5)
    st.iterate(2000.0, 1500.0, 0.1, 0.5, 5)
    T, eps, n = st.auto_parameters()
    # the export byte itself (the saveable mask's union with the frozen export
    # also counts byte 1 as inside, which the statistics do not)
    E = O.unpad_untile(st.state()[2], st.nx, st.ny, st.nz)
    vals = vol[E <= 0].astype(np.float64)
    assert n == vals.size
    assert abs(T - vals.mean()) <= 1e-5 * vals.mean()
    assert abs(eps - vals.std(ddof=1)) <= 1e-4 * vals.std(ddof=1)
