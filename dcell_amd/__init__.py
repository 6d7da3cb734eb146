"""dcell_amd -- MI355X (gfx950) level-set update of DCell.

The product is ``lib/libdcell_lsm.so`` (HIP kernels behind the C ABI of
``include/dcell_lsm.h``) and the C drop-in of ``LevelSetMethod.h``.  This
Python package is the host-side mirror used by tests, the benchmark and the
autograd wrapper of the adjoint.
"""
from . import _lib  # noqa: F401
from .levelset import (GridWindow, advect_sl, curvature, get_velocity, levelset_step,  # noqa: F401
                       precision_of, velocity)

__all__ = ["GridWindow", "advect_sl", "curvature", "get_velocity", "levelset_step", "precision_of", "velocity"]
