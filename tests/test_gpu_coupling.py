"""Two Euler solvers on one grid coupled through their interface ghosts
(coupled_euler_euler_test.cpp:94-165; euler/coupling_conditions.hpp:17-35;
coupling::apply(s1, s2), coupling.hpp:66-73): a Sod tube crossing a cube
that belongs to the second solver, GPU against the oracle."""
import numpy as np
import pytest

import oracle as O
from tests.scenarios import rel_l2

pytestmark = pytest.mark.gpu

SOD = [0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1]


def build(level=4, ti="euler_forward"):
    import hom3_amd as H
    nd = 3
    h = 1.0 / 2 ** level
    L, c = 0.2, [0.5, 0.5, 0.5]
    TI = {"euler_forward": O.TI_EULER_FWD, "ssp_rk3": O.TI_SSP_RK3}[ti]
    # oracle
    og = O.Grid(nd, [0.0] * 3, [1.0] * 3, max_solvers=2).min_level(level)
    p = O.phys(physics=O.PHYS_EULER, nd=nd, ti=TI, cfl=0.2, Minfty=0.2, time_end=0.25)
    o0, o1 = O.Solver(og, p, 0), O.Solver(og, p, 1)
    cube0 = O.geom(O.GEOM_SQUARE, c, [L] * 3)
    o0.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    o0.append_bc(O.geom(O.GEOM_SQUARE, c, [L + h] * 3), O.BC_EULER_COUPLING, [], other=o1)
    o1.append_bc(O.geom(O.GEOM_SQUARE, c, [L - 0.5 * h] * 3, invert=True), O.BC_EULER_COUPLING, [], other=o0)
    o0.set_domain(cube0, +1)
    o1.set_domain(cube0, -1)
    for o in (o0, o1):
        o.set_ic(O.IC_SHOCK_TUBE, SOD)
    o0.initialize()
    o1.initialize()
    # GPU
    g = H.Grid(nd, [0.0] * 3, [1.0] * 3, level, max_solvers=2)
    props = dict(CFL=0.2, Minfty=0.2, timeEnd=0.25)
    s0 = H.Solver(g, physics="euler", time_integration=ti, solver_idx=0, **props)
    s1 = H.Solver(g, physics="euler", time_integration=ti, solver_idx=1, **props)
    cubes = H.make_cube(c, [L] * 3, h, tol=0.5)
    H.append_bcs(s0, H.euler.bc.Neumann())
    s0.append_bc(cubes[1], H.euler.bc.coupling.Euler(s1))
    s1.append_bc(cubes[2], H.euler.bc.coupling.Euler(s0))
    s0.set_initial_domain(cubes[0], +1)
    s1.set_initial_domain(cubes[0], -1)
    for s in (s0, s1):
        s.set_initial_condition(O.IC_SHOCK_TUBE, SOD)
    s0.initialize()
    s1.initialize()
    return (o0, o1), (s0, s1)


def test_euler_euler_coupling():
    """Forward Euler, as in the reference test: the coupling ghosts read the
    other solver's lhs.  (Inside RK stages the condition reads the other
    solver's stage buffer, rhs_tag, which in the reference still holds the
    other solver's previous step, undefined at the first step.)"""
    import hom3_amd as H
    (o0, o1), (s0, s1) = build(ti="euler_forward")
    for o, s in ((o0, s0), (o1, s1)):
        t = s.topology()
        assert s.ncells == o.ncells and s.first_gc == o.first_gc
        for key in ("neighbors", "node_idx", "bc_idx", "x_center", "length"):
            np.testing.assert_array_equal(t[key], o.get(key), err_msg=key)
    assert o1.first_gc == 64 and o1.ncells > o1.first_gc
    nsteps = 6
    for _ in range(nsteps):
        O.couple_same(o0, o1)
        o0.solve()
        o1.solve()
    H.coupling.same_and_step(s0, s1, nsteps)
    for o, s in ((o0, s0), (o1, s1)):
        n = o.first_gc
        err = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], 3)
        assert np.max(err) <= 1e-12, err
        assert abs(s.time() - o.time()) <= 1e-13 * o.time()
