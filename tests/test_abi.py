"""The C ABI library: loads, exports every entry point declared in
include/lsseg/lsseg.h, and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import pytest

from aworkefficientgpualgorithmforlevelsetsegmentation_amd import _build, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lsseg", "lsseg.h")).read()
    return sorted(set(re.findall(r"^(?:lsseg_status|const char\*)\s+(lsseg_\w+)\(", text, re.M)))


def test_header_declarations_match_binding():
    assert _declared() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    path = engine.library_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lsseg_\w+)", out))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    lib = engine.lib()
    for s in _declared():
        assert getattr(lib, s) is not None


def test_library_is_built_for_gfx950():
    out = subprocess.run(["strings", engine.library_path()], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_null_arguments_are_rejected_without_a_device():
    L = engine.lib()
    h = ctypes.c_void_p()
    st = L.lsseg_create(None, None, ctypes.byref(h))
    assert st == engine.ERR_INVALID_ARG
    assert b"null" in L.lsseg_last_error(None)
    assert L.lsseg_destroy(None) == engine.ERR_INVALID_ARG
    assert L.lsseg_iterate(None, None, 1, None) == engine.ERR_INVALID_ARG
    assert L.lsseg_get_mask(None, None) == engine.ERR_INVALID_ARG
    t = engine.Tuning()
    assert L.lsseg_get_tuning(None, ctypes.byref(t)) == engine.ERR_INVALID_ARG
    assert L.lsseg_set_tuning(None, ctypes.byref(t)) == engine.ERR_INVALID_ARG
    assert L.lsseg_autotune(None, None, 10, ctypes.byref(t)) == engine.ERR_INVALID_ARG
    us = ctypes.c_float()
    # the measurement aid checks its arguments before any device call
    assert L.lsseg_probe_random_access(None, 4096, 10, 0, 3, None, ctypes.byref(us)) == engine.ERR_INVALID_ARG
    assert L.lsseg_probe_random_access(ctypes.c_void_p(256), 4096, 10, 5, 3, None, ctypes.byref(us)) == \
        engine.ERR_INVALID_ARG
    assert b"invalid argument" in L.lsseg_last_error(None)


def test_bad_descriptors_are_rejected_without_a_device():
    L = engine.lib()
    buf = (ctypes.c_uint8 * 8)()
    h = ctypes.c_void_p()
    d = engine.VolumeDesc(0, 4, 4, engine.U8, 1.0)
    assert L.lsseg_create(ctypes.byref(d), buf, ctypes.byref(h)) == engine.ERR_INVALID_ARG
    d = engine.VolumeDesc(4, 4, 4, 9, 1.0)
    assert L.lsseg_create(ctypes.byref(d), buf, ctypes.byref(h)) == engine.ERR_UNSUPPORTED
    d = engine.VolumeDesc(2048, 2048, 1024, engine.U8, 1.0)   # 2^32 voxels
    assert L.lsseg_create(ctypes.byref(d), buf, ctypes.byref(h)) == engine.ERR_UNSUPPORTED
    assert b"32-bit" in L.lsseg_last_error(None)


def test_cpp_segmenter_demo_is_built():
    assert os.path.exists(_build.DEMO_BIN) or _build.build_demo()
    res = subprocess.run([_build.DEMO_BIN], capture_output=True, text=True)
    assert res.returncode == 2 and "usage" in res.stderr


@pytest.mark.parametrize("name", ["lsseg.hip", "segmenter.cpp", "synth.c", "segmenter_demo.cpp"])
def test_sources_present(name):
    assert os.path.exists(os.path.join(_build.CSRC, name))
