"""Build recipes for the native parts of the package (in-tree outputs).

* ``liblsseg_hip.so`` -- the HIP engine and its C ABI (``csrc/lsseg.hip``) plus
  the C++ ``lsseg::Segmenter`` host class (``csrc/segmenter.cpp``) and the
  reference's file formats (``csrc/formats.cpp``), compiled by ``hipcc`` for
  gfx950;
* ``_synth.so`` -- the seeded synthetic-volume generators (``csrc/synth.c``),
  compiled by ``gcc``;
* ``segmenter_demo`` -- a small C++ program that drives ``lsseg::Segmenter``
  the way the reference application drives its Segmenter (used by tests).

Nothing here needs a GPU: ``hipcc`` cross-compiles for gfx950.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")

HIP_LIB = os.path.join(PKG_DIR, "liblsseg_hip.so")
SYNTH_LIB = os.path.join(PKG_DIR, "_synth.so")
DEMO_BIN = os.path.join(PKG_DIR, "segmenter_demo")

ARCH = os.environ.get("LSSEG_OFFLOAD_ARCH", "gfx950")

# Device code is compiled without FMA contraction so the update rule rounds
# exactly like the CPU restatement; packed FP32 instructions are disabled.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def _run(cmd, quiet=True):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build command failed ({res.returncode}): {' '.join(cmd)}\n"
                           f"{res.stdout}\n{res.stderr}")
    if not quiet:
        print(res.stdout, res.stderr)


def _atomic_out(path: str) -> str:
    return path + ".tmp"


def build_hip() -> str:
    out = _atomic_out(HIP_LIB)
    srcs = [os.path.join(CSRC, "lsseg.hip"), os.path.join(CSRC, "segmenter.cpp"),
            os.path.join(CSRC, "formats.cpp")]  # + lsseg_device.hpp
    _run([_hipcc(), *HIPCC_FLAGS, "-shared", f"-I{INCLUDE}", "-o", out, *srcs])
    os.replace(out, HIP_LIB)
    return HIP_LIB


def build_synth() -> str:
    out = _atomic_out(SYNTH_LIB)
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall", "-o", out,
          os.path.join(CSRC, "synth.c"), "-lm"])
    os.replace(out, SYNTH_LIB)
    return SYNTH_LIB


def build_demo() -> str:
    out = _atomic_out(DEMO_BIN)
    _run([_hipcc(), "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", out,
          os.path.join(CSRC, "segmenter_demo.cpp"), f"-L{PKG_DIR}", "-llsseg_hip",
          f"-Wl,-rpath,{PKG_DIR}", "-Wl,-rpath,$ORIGIN"])
    os.replace(out, DEMO_BIN)
    return DEMO_BIN


def build_all() -> None:
    build_synth()
    build_hip()
    build_demo()


def synth_library() -> str:
    if not os.path.exists(SYNTH_LIB):
        build_synth()
    return SYNTH_LIB


def hip_library() -> str:
    """Path of the built engine; builds it if absent (hipcc needs no GPU)."""
    if not os.path.exists(HIP_LIB):
        build_hip()
    return HIP_LIB
