"""Python mirror of the reference ``Segmenter`` entry point.

Same method names and meaning as the reference class
(code/gpuseg/code/Segmenter.hpp:40-73) and as the C++ ``lsseg::Segmenter`` in
``include/lsseg/segmenter.hpp``; errors raise ``LssegError``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .engine import MODE_REF, Context

MINIMUM_SEGMENTATION_ITERATIONS = 500  # Config.hpp:45


@dataclass
class Parameters:
    """The Segmenter system of a parameter file (bin/defaults/UnsignedChar.par)."""
    growLevel: float = 150.0
    growWindow: float = 50.0
    timeStep: float = 0.2
    curvatureInfluence: float = 0.3
    numSubSteps: int = 10
    convergenceThreshold: float = 0.0
    mode: int = MODE_REF


class Segmenter:
    def __init__(self, parameters: Parameters | None = None):
        self.parameters = parameters or Parameters()
        self._ctx: Context | None = None
        self._initialized = False
        self._paused = False
        self._num_active = 0
        self.num_iterations = 0
        self._previous = None

    # Segmenter.cpp:110-285
    def LoadVolume(self, volume_zyx: np.ndarray, z_anisotropy: float = 1.0):
        self.UnloadVolume()
        self._ctx = Context(volume_zyx, z_anisotropy)
        return (self._ctx.X, self._ctx.Y, self._ctx.Z)

    def UnloadVolume(self):
        if self._ctx is not None:
            self._ctx.close()
        self._ctx = None
        self._initialized = False
        self._num_active = 0
        self.num_iterations = 0

    @property
    def context(self) -> Context:
        if self._ctx is None:
            raise RuntimeError("no volume loaded")
        return self._ctx

    def _snapshot(self):
        p = self.parameters
        return (p.growLevel, p.growWindow, p.curvatureInfluence)

    # Segmenter.cpp:491-595
    def InitializeSeed(self, seed_xyz, sphere_size: int):
        self.context.seed_sphere(float(seed_xyz[0]), float(seed_xyz[1]), float(seed_xyz[2]), int(sphere_size))
        self._after_init()

    def InitializeLevelSet(self, phi0_zyx: np.ndarray, band: float = 1.0):
        self.context.set_phi(phi0_zyx, band)
        self._after_init()

    def _after_init(self):
        self._num_active = self.context.active_count
        self._initialized = True
        self.num_iterations = 0
        self._previous = self._snapshot()

    def ClearCurrentSegmentation(self):
        self.context.clear(False)
        self._initialized = False
        self._num_active = 0
        self.num_iterations = 0

    def ClearAllSegmentations(self):
        self.context.clear(True)
        self._initialized = False
        self._num_active = 0
        self.num_iterations = 0

    def FreezeCurrentSegmentation(self):
        self.context.freeze()

    def CalculateSegmentationParameters(self):
        """Segmenter.cpp:357-478: growLevel := mean, growWindow := sample standard
        deviation of the source inside the current segmentation."""
        T, eps, _ = self.context.auto_parameters()
        self.parameters.growLevel = T
        self.parameters.growWindow = eps

    def _should_init_active(self) -> bool:
        return self._initialized and self._previous != self._snapshot()  # Segmenter.cpp:1371-1374

    def IsSegmentationInitialized(self) -> bool:
        return self._initialized

    def IsSegmentationFinished(self) -> bool:  # Segmenter.cpp:1053-1056
        return (self._num_active <= self.parameters.convergenceThreshold
                and self.num_iterations > MINIMUM_SEGMENTATION_ITERATIONS)

    def IsSegmentationInProgress(self) -> bool:  # Segmenter.cpp:1358-1361
        return ((self._initialized and not self.IsSegmentationFinished())
                or self._should_init_active()) and not self._paused

    def RequestPlaySegmentation(self):
        self._paused = False

    def RequestPauseSegmentation(self):
        self._paused = True

    # Segmenter.cpp:759-1051
    def RequestUpdateSegmentation(self, time_delta_seconds: float = 0.0):
        if self._ctx is None or not self.IsSegmentationInProgress():
            return None
        p = self.parameters
        steps = int(p.numSubSteps)
        stats = None
        if steps > 0 and self._should_init_active():
            self._ctx.iterate_reinit(p.growLevel, p.growWindow, p.curvatureInfluence, p.timeStep, p.mode)
            self._previous = self._snapshot()
            self.num_iterations = 1
            steps -= 1
        if steps > 0:
            stats = self._ctx.iterate(p.growLevel, p.growWindow, p.curvatureInfluence, p.timeStep, steps, p.mode)
            self.num_iterations += steps
        self._num_active = self._ctx.active_count
        return stats

    def GetActiveElementCount(self) -> int:
        return self._num_active

    # Segmenter.cpp:1216-1337: unpadded int8 {0, 1}
    def GetSaveableSegmentation(self) -> np.ndarray:
        return self.context.mask().astype(np.int8)
