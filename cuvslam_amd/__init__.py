"""MI355X-native sliding-window sparse bundle adjustment (drop-in for cuVSLAM's SchurComplementBundlerGpu)."""
from .sba import (BundleAdjustmentProblem, SchurComplementBundlerGpu, DeviceBatch, SbaError,  # noqa: F401
                  LOSS_STUDENT_T, LOSS_HUBER)
from . import scene  # noqa: F401
