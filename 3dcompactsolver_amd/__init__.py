"""MI355X (gfx950) implementation of the per-stage right-hand-side pipeline
of emathew1/3DCompactSolver: compact derivatives and filter as segmented
line solves, fused stage sweeps, RK4/TVDRK3, behind a C ABI
(include/cs3d.h, lib3dcompact_hip.so).  `cs3d` is the ctypes binding."""
from . import cs3d  # noqa: F401
from .cs3d import (Config, Derivatives, Filter, Solver, config, lib)  # noqa: F401
