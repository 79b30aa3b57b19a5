"""The engine's voxel ids (Morton order inside the reference's 16^3 tiles) and
their neighbours, checked on the host against brute force over every voxel of
several tile geometries, and against the oracle's tiled index.  No GPU: the
helpers are host-and-device functions of csrc/lsseg_device.hpp, compiled here
into a host program."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "aworkefficientgpualgorithmforlevelsetsegmentation_amd", "csrc")


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("hipcc not available")


def test_ids_neighbours_and_tau_conversion_against_brute_force(tmp_path):
    exe = str(tmp_path / "ids_check")
    src = os.path.join(REPO, "tests", "native", "ids_check.cpp")
    res = subprocess.run([_hipcc(), "-O1", "-std=c++17", f"-I{CSRC}", "-o", exe, src], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout + out.stderr


def _spread4(v):
    v = (v | (v << 4)) & 0x0C3
    return (v | (v << 2)) & 0x249


def test_morton_id_is_a_permutation_of_the_oracle_tau_inside_each_tile():
    # restated in numpy from the layout's definition: the id keeps tau's tile
    # and replaces (lz*256 + ly*16 + lx) by the interleaved bits of (lx, ly, lz)
    X, Y, Z = 48, 32, 32
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    t = O.tau(X, Y, x, y, z)
    ident = (t & ~4095) | _spread4(x & 15) | (_spread4(y & 15) << 1) | (_spread4(z & 15) << 2)
    assert np.array_equal(np.sort(ident.ravel()), np.arange(X * Y * Z))
    assert np.array_equal(ident >> 12, t >> 12)
    # a 2x2x2 block shares one 32-byte sector of a float array
    blk = ident[:2, :2, :2].ravel()
    assert np.array_equal(np.sort(blk), np.arange(8))
