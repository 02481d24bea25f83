"""dwost_amd -- Walk-on-Stars (diff_wost) on AMD MI355X.

Scene and solver API of zihay/dwost's ``diff_wost`` package; every query and
walk runs in hand-written gfx950 HIP kernels in ``_build/libdwost.so`` (C ABI:
``include/dwost.h``), also exposed as the PyTorch operators ``torch.ops.dwost.*``
(``dwost_amd.ops``).  There is no CPU fallback: importing the solvers without
the built library raises.
"""

from .shapes import (BC_EXPCOS, BC_LINEAR, BC_PER_VERTEX, BC_QUAD, BoundaryCondition, BoundaryType, Scene2D,
                     Scene3D)
from ._lib import SUB_DIRICHLET, SUB_FULL, SUB_NEUMANN, DwostError
from . import ops  # registers torch.ops.dwost.*
from .solvers import CustomSolver, Solver, WoS, WoS3D, WoSt, WoSt3D, WoStGrad, WoStGrad3D
from .batch import BatchItem, BatchSolver
from .trace import StreamlineTracer

__all__ = [
    "BC_EXPCOS", "BC_LINEAR", "BC_PER_VERTEX", "BC_QUAD", "BoundaryCondition", "BoundaryType", "Scene2D", "Scene3D",
    "SUB_DIRICHLET", "SUB_FULL", "SUB_NEUMANN", "DwostError", "CustomSolver", "Solver", "WoS", "WoS3D", "WoSt", "WoSt3D", "WoStGrad", "WoStGrad3D", "BatchItem", "BatchSolver",
    "StreamlineTracer",
]
