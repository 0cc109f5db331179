"""Workload builders for the benchmark and the tests.

* ``tgv``: 3D compressible Navier-Stokes Taylor-Green vortex on the periodic
  [0, 2 pi]^3 (Re 1600, Mach 0.1, Pr 0.71), AUSM + three-point viscous flux.
* ``vortex``: 2D Euler isentropic vortex on the periodic [0, 10]^2, centre
  (5, 5), AUSM + MUSCL (minmod), SSP-RK3, CFL 0.5.
* ``blast``: 3D Euler blast in the unit cube: base level + refinement of a
  shell of width 2 * band around r = 0.2, 2:1 balanced; inside rho = 1,
  p = 10, outside rho = 0.125, p = 0.1, at rest; a seeded 1% density
  perturbation; zero-gradient (outflow) boundaries; AUSM, SSP-RK3.
* ``channel_cht``: 3D Navier-Stokes channel over a heat-conducting block
  set into its lower wall.  Root [0, 1]^3 at level L (128^3 at L = 7);
  fluid x in [0, 1], y in [0.25, 0.75], z in [0, 0.5] (128 x 64 x 64 cells
  at L = 7), Mach 0.2, inflow at x = 0, outflow at x = 1, adiabatic no-slip
  walls; solid x in [0.375, 0.625], y in [0.125, 0.25], z in [0.125, 0.375]
  (32 x 16 x 32 cells at L = 7) coupled through its top face with
  conductivity ratio kratio, adiabatic elsewhere, with a per-cell source.
  Lengths are in units of the root, so ReInfty = 1000 is Re = 500 on the
  channel height 0.5.
"""
import math

import numpy as np

from . import _lib as L


def tgv(H, level, precision="fp64", ctx=None):
    q = H.Quantities(gamma=1.4, Minfty=0.1)
    g = H.Grid(3, [0.0] * 3, [2 * math.pi] * 3, level, ctx=ctx)
    s = H.Solver(g, physics="cns", time_integration="ssp_rk3", precision=precision, gamma=1.4, CFL=0.5,
                 CFL_viscous=0.5, Minfty=0.1, ReInfty=1600.0, Pr0=0.71)
    H.append_bcs(s, H.Periodic())
    s.set_initial_condition(L.IC_TGV, [q.u_infinity(), q.rho_infinity(), q.p_infinity()])
    s.initialize()
    return g, s


def vortex(H, level=8, precision="fp64", ctx=None):
    g = H.Grid(2, [0.0, 0.0], [10.0, 10.0], level, ctx=ctx)
    s = H.Solver(g, physics="euler", time_integration="ssp_rk3", recon="muscl", limiter="minmod",
                 precision=precision, gamma=1.4, CFL=0.5)
    H.append_bcs(s, H.Periodic())
    s.set_initial_condition(L.IC_ISENTROPIC_VORTEX, [0.0, 5.0, 5.0])
    s.initialize()
    return g, s


BLAST_IC = [0.5, 0.5, 0.5, 0.2, 0.0, 0.0, 0.0, 1.0, 10.0, 0.0, 0.0, 0.0, 0.125, 0.1]


def blast(H, base=6, extra=2, band=0.025, seed=42, amp=0.01, cfl=0.1, ctx=None, precision="fp64"):
    """Returns (grid, solver, refinement counts)."""
    g = H.Grid(3, [0.0] * 3, [1.0] * 3, base, ctx=ctx)
    nref = g.refine_region(base + extra, H.Sphere([0.5, 0.5, 0.5], 0.2), band) if extra else 0
    nbal = g.balance() if extra else 0
    s = H.Solver(g, physics="euler", time_integration="ssp_rk3", CFL=cfl, precision=precision)
    H.append_bcs(s, H.euler.bc.Neumann())
    s.set_initial_condition(L.IC_SPHERE, BLAST_IC)
    s.set_perturbation(seed, amp)
    s.initialize()
    return g, s, (nref, nbal)


CHANNEL = dict(fluid_centre=[0.5, 0.5, 0.25], fluid_dims=[1.0, 0.5, 0.5],
               block_centre=[0.5, 0.1875, 0.25], block_dims=[0.25, 0.125, 0.25])


def channel_cht(H, level=7, kratio=100.0, seed=7, ti="ssp_rk3", Minfty=0.2, ReInfty=1000.0, Pr0=0.71,
                cfl=0.4, heat_cfl=0.3, ctx=None, kernel="auto", precision="fp64"):
    """Returns (grid, cns, heat, source).  kernel: the fluid solver's kernel
    family; precision: both solvers' state precision (the interface
    temperatures stay FP64)."""
    nd = 3
    g = H.Grid(nd, [0.0] * 3, [1.0] * 3, level, max_solvers=2, ctx=ctx)
    h = 1.0 / 2 ** level
    q = H.Quantities(gamma=1.4, Minfty=Minfty)
    pv = [q.u_infinity(), 0.0, 0.0, q.rho_infinity(), q.p_infinity()]
    cns = H.Solver(g, physics="cns", time_integration=ti, solver_idx=0, CFL=cfl, CFL_viscous=cfl,
                   Minfty=Minfty, ReInfty=ReInfty, Pr0=Pr0, timeEnd=1e30, kernel=kernel, precision=precision)
    heat = H.Solver(g, physics="heat", flux="three_point", time_integration=ti, solver_idx=1, CFL=heat_cfl,
                    timeEnd=1e30, precision=precision)
    C = CHANNEL
    fluid = H.invert(H.Square(C["fluid_centre"], C["fluid_dims"]))
    block = H.invert(H.Square(C["block_centre"], C["block_dims"]))
    cube1 = H.Square(C["block_centre"], [d + h for d in C["block_dims"]])
    cube2 = H.invert(H.Square(C["block_centre"], [d - 0.6 * h for d in C["block_dims"]]))
    cube = H.cube_boundaries(g)
    y0, y1, z1 = 0.25, 0.75, 0.5
    # fluid: inflow, outflow, the conjugate interface, then the walls (the
    # first condition whose geometry claims a missing neighbour owns it)
    cns.append_bc(cube[0], H.cns.bc.external.Inflow(pv))
    cns.append_bc(cube[1], H.cns.bc.external.Outflow(pv))
    cns.append_bc(cube1, H.cns.bc.coupling.Heat(heat, kratio))
    cns.append_bc(H.Edge([0.5, y0, 0.25], [0.0, 1.0, 0.0]), H.cns.bc.wall.AdiabaticNoSlip())
    cns.append_bc(H.Edge([0.5, y1, 0.25], [0.0, -1.0, 0.0]), H.cns.bc.wall.AdiabaticNoSlip())
    cns.append_bc(cube[4], H.cns.bc.wall.AdiabaticNoSlip())
    cns.append_bc(H.Edge([0.5, 0.5, z1], [0.0, 0.0, -1.0]), H.cns.bc.wall.AdiabaticNoSlip())
    # solid: the top face couples to the fluid, the other faces are adiabatic
    heat.append_bc(H.Edge([0.5, y0, 0.25], [0.0, -1.0, 0.0]), H.heat.bc.coupling.CNS(cns, kratio))
    heat.append_bc(cube2, H.heat.bc.Neumann(0.0))
    cns.set_initial_domain(fluid, +1)
    heat.set_initial_domain(block, +1)
    cns.set_initial_condition(L.IC_CONSTANT_PV, pv)
    heat.set_initial_condition(L.IC_HEAT_CONSTANT, [q.T_infinity()])
    heat.initialize()
    cns.initialize()
    src = np.random.default_rng(seed).uniform(0.0, 1.0, heat.first_gc)
    heat.set_source(src)
    return g, cns, heat, src
