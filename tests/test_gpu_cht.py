"""Conjugate heat transfer: a CNS solver and a heat solver sharing one grid,
coupled through the interface ghost cells (coupled_cns_heat_test.cpp:160-250,
cns/coupling_conditions.hpp:21-43, heat/coupling_conditions.hpp:18-38,
coupling.hpp:47-63), GPU against the oracle."""
import numpy as np
import pytest

import oracle as O
from tests.scenarios import rel_l2

pytestmark = pytest.mark.gpu


def cht_setup_oracle(Om, nd=2, level=5, half=4.0, kratio=1.0, source=None, ti="euler_forward", time_end=(10.0, 10.0)):
    """The oracle side of cht_setup, with the oracle front-end Om (the FP64
    build or its long-double twin); time_end: (fluid, solid) "timeEnd"."""
    import hom3_amd as H
    xmin, xmax = [-half] * nd, [half] * nd
    h = 2 * half / 2 ** level
    centre = [-1.0] + [0.0] * (nd - 1)
    dims = [1.0] * nd
    q = H.Quantities(gamma=1.4, Minfty=0.2)
    pv = [0.0] * (nd + 2)
    pv[0] = q.u_infinity()
    pv[nd] = q.rho_infinity()
    pv[nd + 1] = q.p_infinity()
    TI = {"euler_forward": Om.TI_EULER_FWD, "ssp_rk3": Om.TI_SSP_RK3}[ti]
    og = Om.Grid(nd, xmin, xmax, max_solvers=2).min_level(level)
    ocns = Om.Solver(og, Om.phys(physics=Om.PHYS_CNS, nd=nd, ti=TI, cfl=0.4, cfl_viscous=0.4, Minfty=0.2,
                                 ReInfty=100.0, Pr0=0.72, time_end=time_end[0]), 0)
    oheat = Om.Solver(og, Om.phys(physics=Om.PHYS_HEAT, nd=nd, flux=Om.FLUX_THREE_POINT, ti=TI, cfl=0.5,
                                  time_end=time_end[1]), 1)
    geos = Om.cube_boundaries(nd, xmin, xmax)
    ocns.append_bc(geos[0], Om.BC_CNS_INFLOW, pv)
    ocns.append_bc(geos[1], Om.BC_CNS_OUTFLOW, pv)
    for gg in geos[2:]:
        ocns.append_bc(gg, Om.BC_EULER_NEUMANN, [])
    og0 = Om.geom(Om.GEOM_SQUARE, centre, dims)
    og1 = Om.geom(Om.GEOM_SQUARE, centre, [d + h for d in dims])
    og2 = Om.geom(Om.GEOM_SQUARE, centre, [d - 0.6 * h for d in dims], invert=True)
    ocns.append_bc(og1, Om.BC_CHT_FLUID, [], other=oheat, kratio=kratio)
    oheat.append_bc(og2, Om.BC_CHT_SOLID, [], other=ocns, kratio=kratio)
    ocns.set_domain(og0, +1)
    oheat.set_domain(og0, -1)
    ocns.set_ic(Om.IC_CONSTANT_PV, pv)
    oheat.set_ic(Om.IC_HEAT_CONSTANT, [q.T_infinity() * 1.05])
    oheat.initialize()
    ocns.initialize()
    if source is not None:
        oheat.set_source(source(oheat.first_gc))
    return ocns, oheat


def cht_setup(nd=2, level=5, half=4.0, kratio=1.0, source=None, ti="euler_forward", time_end=(10.0, 10.0)):
    """Flow past a heat-conducting block centred at (-1, 0[, 0]), side 1."""
    import hom3_amd as H
    xmin, xmax = [-half] * nd, [half] * nd
    h = 2 * half / 2 ** level
    centre = [-1.0] + [0.0] * (nd - 1)
    dims = [1.0] * nd
    cube0, cube1, cube2 = H.make_cube(centre, dims, h)
    q = H.Quantities(gamma=1.4, Minfty=0.2)
    pv = [0.0] * (nd + 2)
    pv[0] = q.u_infinity()
    pv[nd] = q.rho_infinity()
    pv[nd + 1] = q.p_infinity()
    cns_p = dict(CFL=0.4, CFL_viscous=0.4, Minfty=0.2, ReInfty=100.0, Pr0=0.72, timeEnd=time_end[0])
    heat_p = dict(CFL=0.5, timeEnd=time_end[1])

    # ---- oracle
    ocns, oheat = cht_setup_oracle(O, nd, level, half, kratio, source, ti, time_end)

    # ---- GPU
    g = H.Grid(nd, xmin, xmax, level, max_solvers=2)
    cns = H.Solver(g, physics="cns", time_integration=ti, solver_idx=0, **cns_p)
    heat = H.Solver(g, physics="heat", flux="three_point", time_integration=ti, solver_idx=1, **heat_p)
    geos_h = H.cube_boundaries(g)
    cns.append_bc(geos_h[0], H.cns.bc.external.Inflow(pv))
    cns.append_bc(geos_h[1], H.cns.bc.external.Outflow(pv))
    for gg in geos_h[2:]:
        cns.append_bc(gg, H.cns.bc.Neumann())
    cns.append_bc(cube1, H.cns.bc.coupling.Heat(heat, kratio))
    heat.append_bc(cube2, H.heat.bc.coupling.CNS(cns, kratio))
    cns.set_initial_domain(cube0, +1)
    heat.set_initial_domain(cube0, -1)
    cns.set_initial_condition(0, pv)
    heat.set_initial_condition(6, [q.T_infinity() * 1.05])
    heat.initialize()
    cns.initialize()
    if source is not None:
        heat.set_source(source(heat.first_gc))
    for o, s in ((ocns, cns), (oheat, heat)):
        s.set_lhs(o.get("lhs"))
        s.reset_time(o.time(), o.dt(), o.step())
    return (ocns, oheat), (cns, heat)


def _topo_eq(o, s):
    t = s.topology()
    assert s.ncells == o.ncells and s.first_gc == o.first_gc
    for key in ("neighbors", "node_idx", "bc_idx", "x_center", "length"):
        np.testing.assert_array_equal(t[key], o.get(key), err_msg=key)


@pytest.mark.parametrize("nd,level", [(2, 5), (3, 4)])
def test_cht_topology_and_partners(nd, level):
    (oc, oh), (c, h) = cht_setup(nd, level)
    _topo_eq(oc, c)
    _topo_eq(oh, h)
    assert oh.ncells > oh.first_gc and oc.ncells > oc.first_gc


@pytest.mark.parametrize("nd,level,kratio,ti", [(2, 5, 1.0, "euler_forward"), (3, 4, 1.0, "euler_forward"),
                                                (2, 5, 100.0, "ssp_rk3")])
def test_cht_coupled_steps(nd, level, kratio, ti):
    src = None
    if kratio != 1.0:
        def src(n):   # the same seeded field for the oracle and the GPU
            return np.random.default_rng(5).uniform(0, 1, n)
    (oc, oh), (c, h) = cht_setup(nd, level, kratio=kratio, source=src, ti=ti)
    import hom3_amd as H
    nsteps = 6
    for _ in range(nsteps):
        O.couple_cns_heat(0.01, oc, oh)
        oc.solve()
        oh.solve()
    H.coupling.apply_and_step(0.01, c, h, nsteps)
    for o, s in ((oc, c), (oh, h)):
        n = o.first_gc
        err = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], nd if o.nvars > 1 else None)
        assert np.max(err) <= 1e-12, err
        assert abs(s.time() - o.time()) <= 1e-13 * max(1, o.time())
        assert abs(s.dt() - o.dt()) <= 1e-13 * o.dt()
