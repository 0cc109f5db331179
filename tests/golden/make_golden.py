"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

The fixtures pin configurations the reference cannot run (periodic
conditions, MUSCL, SSP-RK3, Taylor-Green) and reference-mode runs, as small
FP64 arrays: the initial state, the state after N steps, dt and time.
Regenerate with:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def vortex_muscl():
    g = O.Grid(2, [-10, -10], [0, 0]).min_level(4)
    s = O.Solver(g, O.phys(nd=2, ti=O.TI_SSP_RK3, cfl=0.5, recon=O.RECON_MUSCL, limiter=O.LIM_MINMOD))
    s.append_cube_bcs([O.BC_PERIODIC] * 4)
    s.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0])
    return g, s, 5


def tgv():
    g = O.Grid(3, [0] * 3, [2 * math.pi] * 3).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_CNS, nd=3, ti=O.TI_SSP_RK3, cfl=0.5, cfl_viscous=0.5, Minfty=0.1,
                           ReInfty=1600.0, Pr0=0.71))
    s.append_cube_bcs([O.BC_PERIODIC] * 6)
    q = s.quantities()
    s.set_ic(O.IC_TGV, [q["u_inf"], q["rho_inf"], q["p_inf"]])
    return g, s, 2


def sod_reference():
    g = O.Grid(3, [0] * 3, [1] * 3).min_level(3)
    s = O.Solver(g, O.phys(nd=3, ti=O.TI_EULER_FWD, cfl=0.2))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    s.set_ic(O.IC_SHOCK_TUBE, [0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1])
    return g, s, 4


def cns_walls_reference():
    g = O.Grid(2, [0, 0], [1, 1]).min_level(4)
    s = O.Solver(g, O.phys(physics=O.PHYS_CNS, nd=2, ti=O.TI_EULER_FWD, cfl=0.4, cfl_viscous=0.4, Minfty=0.1,
                           ReInfty=1000.0))
    s.append_cube_bcs([O.BC_ADIABATIC_NOSLIP] * 4)
    q = s.quantities()
    s.set_ic(O.IC_SHOCK_TUBE, [0, 0, 0.3, q["rho_inf"], 0.75 * q["u_inf"], q["p_inf"], 0.125 * q["rho_inf"], 0.0,
                               0.1 * q["p_inf"]])
    return g, s, 4


CASES = {"vortex_muscl_rk3_16x16": vortex_muscl, "tgv_cns_rk3_8x8x8": tgv,
         "sod_euler_fe_8x8x8": sod_reference, "cns_walls_fe_16x16": cns_walls_reference}


def make(name):
    g, s, nsteps = CASES[name]()
    s.initialize()
    q0 = s.get("lhs")[:, : s.first_gc].copy()
    dt0 = s.dt()
    s.solve(nsteps)
    return dict(q0=q0, q1=s.get("lhs")[:, : s.first_gc].copy(), dt0=np.float64(dt0), dt1=np.float64(s.dt()),
                time=np.float64(s.time()), nsteps=np.int64(nsteps))


if __name__ == "__main__":
    for name in CASES:
        np.savez(os.path.join(HERE, name + ".npz"), **make(name))
        print("wrote", name)
