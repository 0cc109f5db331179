"""hom3_amd — MI355X-native finite-volume hot path of gnzlbg/hom3.

The compute path is libhom3g.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/hom3g.h).  This package is the host-side mirror of the
reference's solver interface (see hom3_amd.fv) plus the build script.
"""
from . import _lib  # noqa: F401
from .fv import (Context, Grid, Solver, Quantities, Condition, Periodic, Edge, Square, Sphere,  # noqa: F401
                 invert, make_cube, cube_boundaries, append_bcs, coupling, run_solver,
                 solver_finished, solution_diverged, SolutionDiverged, default_context, euler, cns,
                 heat, HostCondition, GhostSet)
from .batch import Batch, four_quadrant_instances  # noqa: F401
from .steplog import StepLog  # noqa: F401

__all__ = ["Context", "Grid", "Solver", "Quantities", "Condition", "Periodic", "Edge", "Square",
           "Sphere", "invert", "make_cube", "cube_boundaries", "append_bcs", "coupling", "run_solver",
           "euler", "cns", "heat", "Batch", "four_quadrant_instances", "StepLog", "HostCondition", "GhostSet",
           "SolutionDiverged"]
