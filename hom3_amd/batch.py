"""Batched first-order Euler instances on one device (hom3g_batch_*).

Each instance is an ordinary built and initialised ``fv.Solver`` (its own
grid, 2:1 refinement, dt, time and step); a ``Batch`` copies their state and
advances all of them with one persistent kernel per RK stage, one step per
hipGraph launch.  ``four_quadrant_instances`` draws the seeded family of
uneven 2D four-quadrant Riemann problems used by the batched workload.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib

PRECISION = {"fp64": L.FP64, "fp32": L.FP32_STATE}


class Batch:
    def __init__(self, solvers, precision="fp64"):
        self.solvers = list(solvers)
        self.ctx = self.solvers[0].ctx
        self.nvars = self.solvers[0].nvars
        self.n_int = [s.first_gc for s in self.solvers]
        ids = np.ascontiguousarray([s.id for s in self.solvers], np.int32)
        bid = C.c_int()
        check(lib().hom3g_batch_create(self.ctx.h, ids.ctypes.data_as(C.POINTER(C.c_int32)), ids.size,
                                       PRECISION[precision], C.byref(bid)))
        self.id = bid.value
        self.precision = precision

    def set_config(self, block, tile, grid):
        check(lib().hom3g_batch_set_config(self.ctx.h, self.id, block, tile, grid))

    def config(self):
        b, t, g, n = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        check(lib().hom3g_batch_get_config(self.ctx.h, self.id, C.byref(b), C.byref(t), C.byref(g), C.byref(n)))
        return dict(block=b.value, tile=t.value, grid=g.value, items=n.value)

    def autotune(self):
        """Pick block/tile/grid by timing stage 1; returns the choice and its time."""
        b, t, g, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        check(lib().hom3g_batch_autotune(self.ctx.h, self.id, C.byref(b), C.byref(t), C.byref(g), C.byref(ms)))
        return dict(block=b.value, tile=t.value, grid=g.value, stage_ms=ms.value)

    def solve(self, nsteps=1, graph=True):
        check(lib().hom3g_batch_solve(self.ctx.h, self.id, nsteps, 1 if graph else 0))

    def time_solve(self, nsteps, graph=True):
        """ms per step between two events on the library stream."""
        ms = C.c_float()
        check(lib().hom3g_batch_time_solve(self.ctx.h, self.id, nsteps, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def counts(self):
        n, c = C.c_int(), C.c_int64()
        check(lib().hom3g_batch_counts(self.ctx.h, self.id, C.byref(n), C.byref(c)))
        return n.value, c.value

    def state(self, i):
        n = self.n_int[i]
        out = np.empty((self.nvars, n))
        check(lib().hom3g_batch_download_state(self.ctx.h, self.id, i, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def time(self, i):
        t, dt, st = C.c_double(), C.c_double(), C.c_int64()
        check(lib().hom3g_batch_time(self.ctx.h, self.id, i, C.byref(t), C.byref(dt), C.byref(st)))
        return t.value, dt.value, st.value

    def close(self):
        if self.id is not None:
            check(lib().hom3g_batch_destroy(self.ctx.h, self.id))
            self.id = None


def four_quadrant_instances(seed, n, levels=(6, 7, 8, 9), max_extra=2):
    """Seeded parameters of n 2D four-quadrant Riemann instances on [0,1]^2:
    base level uniform over ``levels`` (log-uniform sides 2^L), 0..max_extra
    refinement levels inside a disc with centre U[0,1]^2 and radius
    U[0.1,0.3], and per quadrant rho, p ~ U[0.1,2], u, v ~ U[-1,1]
    (quadrants split at (0.5, 0.5))."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        level = int(rng.choice(levels))
        extra = int(rng.integers(0, max_extra + 1))
        centre = rng.uniform(0.0, 1.0, 2).tolist()
        radius = float(rng.uniform(0.1, 0.3))
        pv = []
        for _q in range(4):
            u, v = rng.uniform(-1.0, 1.0, 2)
            rho, p = rng.uniform(0.1, 2.0, 2)
            pv += [float(u), float(v), float(rho), float(p)]
        out.append(dict(level=level, extra=extra, centre=centre, radius=radius,
                        ic=[0.5, 0.5] + pv))
    return out


def build_instance(H, prm, ti="ssp_rk3", cfl=0.4, ctx=None):
    """One instance as an fv.Solver: MinLevel(level), refinement of the disc
    (a sphere of radius R/2 with band R/2 covers the disc), 2:1 balance,
    zero-gradient Neumann boundaries, four-quadrant initial condition."""
    g = H.Grid(2, [0.0, 0.0], [1.0, 1.0], prm["level"], ctx=ctx)
    if prm["extra"] > 0:
        half = 0.5 * prm["radius"]
        g.refine_region(prm["level"] + prm["extra"], H.Sphere(prm["centre"], half), half)
        g.balance()
    s = H.Solver(g, physics="euler", time_integration=ti, CFL=cfl)
    H.append_bcs(s, H.euler.bc.Neumann())
    s.set_initial_condition(L.IC_FOUR_QUADRANT, prm["ic"])
    s.initialize()
    return s
