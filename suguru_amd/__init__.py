"""suguru_amd -- MI355X (gfx950) immersed-boundary fluid-membrane transfer.

The hot path of the reference capsule solver (interp_spread.f90, kernel.f90): interpolation
of the staggered MAC velocity to capsule points and spreading of membrane forces to the grid,
as hand-written FP64 HIP kernels behind a C ABI (include/ibx.h, libibx.so).
"""
from .interp_spread import IBTransfer, Batch, IbxError, kernel  # noqa: F401
from . import workload  # noqa: F401

__all__ = ["IBTransfer", "Batch", "IbxError", "kernel", "workload"]
