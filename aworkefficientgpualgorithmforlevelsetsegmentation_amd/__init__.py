"""MI355X-native work-efficient sparse level-set segmentation.

The engine is ``liblsseg_hip.so`` (HIP kernels for gfx950 behind the C ABI in
``include/lsseg/lsseg.h``); ``engine.Context`` binds that ABI and
``segmenter.Segmenter`` mirrors the reference's ``Segmenter`` entry point.
``synth`` generates the seeded benchmark volumes.
"""
from .engine import Context, LssegError, MODE_CLEAN, MODE_REF  # noqa: F401
from .segmenter import Parameters, Segmenter  # noqa: F401

__all__ = ["Context", "LssegError", "MODE_REF", "MODE_CLEAN", "Segmenter", "Parameters"]
