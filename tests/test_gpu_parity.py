"""Parity of the HIP path (libhom3g through the C ABI) against the CPU oracle.

Topology (cell order, neighbour ids, ghost map, coordinates) must match
bit-exactly; states after N steps must agree to 1e-12 relative L2 per
conserved variable (the FP64 parity bar).
"""
import numpy as np
import pytest

import oracle as O
from tests.scenarios import build_gpu, build_oracle, cube_bcs, rel_l2, scenario

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def H():
    import hom3_amd
    return hom3_amd


# --------------------------------------------------------------------------- grid
@pytest.mark.parametrize("nd,level", [(2, 3), (3, 2), (2, 5), (3, 4)])
def test_grid_tree_neighbors_geometry(H, nd, level):
    og = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(level)
    hg = H.Grid(nd, [0.0] * nd, [1.0] * nd, level)
    assert hg.size() == og.size()
    p1, f1 = og.tree()
    p2, f2 = hg.tree()
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(og.all_neighbors(), hg.neighbors())
    x, ln = hg.geometry()
    for n in range(0, og.size(), max(1, og.size() // 97)):
        assert np.array_equal(x[n], og.cell_coordinates(n))
        assert ln[n] == og.cell_length(n)


def _topology_equal(os_, hs):
    t = hs.topology()
    assert hs.ncells == os_.ncells and hs.first_gc == os_.first_gc
    for key in ("neighbors", "node_idx", "bc_idx", "x_center", "length"):
        np.testing.assert_array_equal(t[key], os_.get(key), err_msg=key)
    d_o = os_.get("distances")
    np.testing.assert_array_equal(np.isnan(t["distances"]), np.isnan(d_o))
    np.testing.assert_array_equal(np.nan_to_num(t["distances"]), np.nan_to_num(d_o))


def _cube_cfg(nd, x0, length, h):
    centre = [x0] + [0.0] * (nd - 1)
    return centre, [length] * nd


TOPO_SCENARIOS = {
    "euler2d_neumann": lambda: _with(scenario(2, 4), "neumann"),
    "euler3d_periodic": lambda: _with(scenario(3, 3), "periodic"),
    "cns2d_walls": lambda: _with(scenario(2, 4, "cns"), "wall"),
    "cns2d_cube": lambda: _flow_past_cube(2, 5),
    "cns3d_cube": lambda: _flow_past_cube(3, 4),
}


def _with(sc, kind):
    k = {"neumann": O.BC_EULER_NEUMANN, "periodic": O.BC_PERIODIC, "wall": O.BC_ADIABATIC_NOSLIP}[kind]
    sc["bcs"] = cube_bcs(sc, k)
    sc["ic"] = (O.IC_SHOCK_TUBE, [0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1])
    return sc


def _freestream(sc):
    import hom3_amd as H
    q = H.Quantities(gamma=sc["props"]["gamma"], Minfty=sc["props"]["Minfty"])
    nd = sc["nd"]
    pv = [0.0] * (nd + 2)
    pv[0] = q.u_infinity()
    pv[nd] = q.rho_infinity()
    pv[nd + 1] = q.p_infinity()
    return pv


def _flow_past_cube(nd, level):
    """cns_fv_solver_test.cpp:212-282 (flow past a cube) on a [-7,7]^nd root."""
    sc = scenario(nd, level, "cns", xmin=[-7.0] * nd, xmax=[7.0] * nd, CFL=0.4, CFL_viscous=0.4,
                  ReInfty=200.0)
    h = 14.0 / 2 ** level
    centre, dims = [-3.0] + [0.0] * (nd - 1), [1.0] * nd
    pv = _freestream(sc)
    bcs = [("cube", 0, O.BC_CNS_INFLOW, pv), ("cube", 1, O.BC_CNS_OUTFLOW, pv)]
    for i in range(2, 2 * nd):
        bcs.append(("cube", i, O.BC_EULER_NEUMANN, []))
    bcs.append((("square", centre, [d + h for d in dims]), O.BC_ADIABATIC_NOSLIP, []))
    sc["bcs"] = bcs
    sc["domain"] = (("square", centre, dims), +1)
    # density bump upstream of the cube (cns_fv_solver_test.cpp:239-251) as a sphere state
    sc["ic"] = (O.IC_SPHERE, [-4.5] + [0.4] * (nd - 1) + [0.45] + _bump(pv, nd) + pv)
    return sc


def _bump(pv, nd):
    b = list(pv)
    b[nd] = pv[nd] * 0.9
    return b


@pytest.mark.parametrize("name", list(TOPO_SCENARIOS))
def test_topology_bit_exact(H, name):
    sc = TOPO_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    _topology_equal(os_, hs)


# --------------------------------------------------------------------------- stepping
def _sync_state(os_, hs):
    """Start both from the oracle's initial state and time."""
    hs.set_lhs(os_.get("lhs"))
    hs.reset_time(os_.time(), os_.dt(), os_.step())


def _compare(os_, hs, tol=TOL):
    n = os_.first_gc
    qo = os_.get("lhs")[:, :n]
    qh = hs.lhs[:, :n]
    err = rel_l2(qh, qo, os_.nd)
    assert np.all(np.isfinite(qh)), "non-finite GPU state"
    assert np.max(err) <= tol, f"rel L2 per variable {err}"
    assert abs(hs.time() - os_.time()) <= 1e-13 * max(1.0, abs(os_.time()))
    assert hs.step() == os_.step()
    return err


STEP_SCENARIOS = {
    "euler2d_sod_fe": lambda: _with(scenario(2, 5, ti="euler_forward"), "neumann"),
    "euler3d_sod_rk2": lambda: _with(scenario(3, 3, ti="runge_kutta_2"), "neumann"),
    "euler2d_sod_rk3_llf": lambda: _with(scenario(2, 5, ti="ssp_rk3", flux="llf"), "neumann"),
    "euler2d_vortex_periodic_muscl": lambda: _vortex(recon="muscl"),
    "euler2d_vortex_bc": lambda: _vortex_bc(),
    "cns2d_sod_walls_fe": lambda: _cns_sod(2, 5, "wall", "euler_forward"),
    "cns3d_sod_slipwalls_rk3": lambda: _cns_sod(3, 3, "slip", "ssp_rk3"),
    "cns2d_cube_fe": lambda: _flow_past_cube(2, 5),
    "cns3d_tgv_rk3_generic": lambda: _tgv(3),
    "cns3d_tgv_rk3_brick": lambda: _tgv(3, 4),
    "euler3d_periodic_rk3_brick": lambda: _periodic_euler3d(),
    "euler2d_neumann_rk3_brick": lambda: _with(scenario(2, 5, ti="ssp_rk3"), "neumann"),
    "heat2d_dirichlet": lambda: _heat(2, 5),
    "heat3d_neumann_rk2": lambda: _heat(3, 3, neumann=True),
    "euler2d_dirichlet_rk3": lambda: _euler_dirichlet(2, 5, "ssp_rk3"),
    "euler3d_dirichlet_fe": lambda: _euler_dirichlet(3, 3, "euler_forward"),
    "cns2d_isothermal_walls_rk2": lambda: _cns_isothermal(2, 5, "runge_kutta_2"),
    "cns3d_isothermal_walls_rk3": lambda: _cns_isothermal(3, 3, "ssp_rk3"),
    # several bricks per solver on tiled bricks (Euler and CNS, 3D)
    "euler3d_dirichlet_rk3_8bricks": lambda: _euler_dirichlet(3, 4, "ssp_rk3"),
    "cns3d_isothermal_walls_rk3_8bricks": lambda: _cns_isothermal(3, 4, "ssp_rk3"),
}


def _euler_dirichlet(nd, level, ti):
    """euler/boundary_conditions.hpp:42-60: constant conserved surface values,
    the outer state perturbed by a different 2% on each face, around a dense
    high-pressure sphere."""
    sc = scenario(nd, level, ti=ti, CFL=0.3)
    rng = np.random.default_rng(nd)
    out = [0.1, 0.05, 0.0][:nd] + [1.0, 1.0]
    inside = [0.1, 0.05, 0.0][:nd] + [1.3, 1.5]

    def cv(pv):
        rho, p = pv[nd], pv[nd + 1]
        return [rho * u for u in pv[:nd]] + [rho, p / 0.4 + 0.5 * rho * sum(u * u for u in pv[:nd])]
    sc["bcs"] = [("cube", i, O.BC_EULER_DIRICHLET, cv([x * (1 + 0.02 * rng.uniform(-1, 1)) for x in out]))
                 for i in range(2 * nd)]
    sc["ic"] = (O.IC_SPHERE, [0.5] * nd + [0.25] + inside + out)
    return sc


def _cns_isothermal(nd, level, ti):
    """cns/boundary_conditions.hpp:199-238: isothermal no-slip walls 10%
    hotter than the free stream around a Sod tube."""
    sc = scenario(nd, level, "cns", ti=ti, CFL=0.4, CFL_viscous=0.4, ReInfty=1000.0)
    import hom3_amd as H
    q = H.Quantities(gamma=1.4, Minfty=0.1)
    sc["bcs"] = cube_bcs(sc, O.BC_ISOTHERMAL_NOSLIP, [1.1 * q.T_infinity()])
    sc["ic"] = (O.IC_SHOCK_TUBE, [0, 0, 0.3, q.rho_infinity(), 0.75 * q.u_infinity(), q.p_infinity(),
                                  0.125 * q.rho_infinity(), 0.0, 0.1 * q.p_infinity()])
    return sc


def _vortex(recon="none", level=5):
    """Isentropic vortex on the periodic domain [-10, 0]^2 (reference centre (-5,-5))."""
    sc = scenario(2, level, xmin=[-10.0, -10.0], xmax=[0.0, 0.0], ti="ssp_rk3", recon=recon,
                  limiter="minmod", CFL=0.5)
    sc["bcs"] = cube_bcs(sc, O.BC_PERIODIC)
    sc["ic"] = (O.IC_ISENTROPIC_VORTEX, [0.0])
    return sc


def _vortex_bc():
    """euler_fv_solver_test.cpp:258-286: vortex with the analytic BC on [-8,8]^2."""
    sc = scenario(2, 5, xmin=[-8.0, -8.0], xmax=[8.0, 8.0], ti="runge_kutta_2", CFL=0.2)
    sc["bcs"] = cube_bcs(sc, O.BC_ISENTROPIC_VORTEX)
    sc["ic"] = (O.IC_ISENTROPIC_VORTEX, [0.0])
    return sc


def _cns_sod(nd, level, wall, ti):
    sc = scenario(nd, level, "cns", ti=ti, CFL=0.4, CFL_viscous=0.4, ReInfty=1000.0)
    import hom3_amd as H
    q = H.Quantities(gamma=1.4, Minfty=0.1)
    k = O.BC_ADIABATIC_NOSLIP if wall == "wall" else O.BC_ADIABATIC_SLIP
    sc["bcs"] = cube_bcs(sc, k)
    sc["ic"] = (O.IC_SHOCK_TUBE, [0, 0, 0.3, q.rho_infinity(), 0.75 * q.u_infinity(), q.p_infinity(),
                                  0.125 * q.rho_infinity(), 0.0, 0.1 * q.p_infinity()])
    return sc


def _tgv(nd=3, level=3):
    import math
    sc = scenario(nd, level, "cns", xmin=[0.0] * nd, xmax=[2 * math.pi] * nd, ti="ssp_rk3", CFL=0.5,
                  CFL_viscous=0.5, Minfty=0.1, ReInfty=1600.0, Pr0=0.71)
    import hom3_amd as H
    q = H.Quantities(gamma=1.4, Minfty=0.1)
    sc["bcs"] = cube_bcs(sc, O.BC_PERIODIC)
    sc["ic"] = (O.IC_TGV, [q.u_infinity(), q.rho_infinity(), q.p_infinity()])
    return sc


def _periodic_euler3d():
    sc = scenario(3, 4, ti="ssp_rk3", CFL=0.5)
    sc["bcs"] = cube_bcs(sc, O.BC_PERIODIC)
    sc["ic"] = (O.IC_SPHERE, [0.5, 0.5, 0.5, 0.2, 0.1, -0.2, 0.05, 1.0, 1.0, 0.0, 0.3, 0.1, 0.125, 0.1])
    return sc


def _heat(nd, level, neumann=False):
    sc = scenario(nd, level, "heat", ti="runge_kutta_2" if neumann else "euler_forward", CFL=0.4)
    if neumann:
        sc["bcs"] = cube_bcs(sc, O.BC_HEAT_NEUMANN, [0.5])
    else:
        sc["bcs"] = [("cube", i, O.BC_HEAT_DIRICHLET, [1.0 + 0.1 * i]) for i in range(2 * nd)]
    sc["ic"] = (O.IC_HEAT_SINE, [1.0, 0.3, 3.0, 0.1])
    return sc


# every scenario with the kernel picked at build ("auto": bricks, tiled
# bricks where the cells fall into whole bricks, else table-driven) and with
# the table-driven kernel forced
_PARITY_CASES = [(n, "auto") for n in STEP_SCENARIOS if not n.endswith("generic")] + \
    [(n, "generic") for n in STEP_SCENARIOS if not n.startswith("heat")]


@pytest.mark.parametrize("name,kernel", _PARITY_CASES)
def test_step_parity(H, name, kernel):
    sc = STEP_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc, kernel=kernel)
    if name.endswith("brick") and kernel == "auto":
        assert hs.kernel() == "brick"
    _topology_equal(os_, hs)
    _sync_state(os_, hs)
    nsteps = 6
    os_.solve(nsteps)
    hs.solve(nsteps)
    err = _compare(os_, hs)
    print(name, "rel L2", err)


@pytest.mark.parametrize("name", ["euler2d_sod_fe", "cns3d_tgv_rk3_generic", "euler2d_vortex_periodic_muscl"])
def test_initial_condition(H, name):
    sc = STEP_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    n = os_.first_gc
    err = rel_l2(hs.lhs[:, :n], os_.get("lhs")[:, :n], os_.nd)
    assert np.max(err) < 1e-13, err
    assert abs(hs.dt() - os_.dt()) <= 1e-13 * os_.dt()


def test_set_dt_clamps_to_time_end(H):
    """set_dt uses the previous dt to decide the final step (solver.hpp:541-551)."""
    sc = _with(scenario(2, 4, ti="euler_forward", timeEnd=0.02), "neumann")
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    _sync_state(os_, hs)
    for _ in range(6):
        os_.solve()
        hs.solve()
        assert abs(hs.dt() - os_.dt()) <= 1e-14 * abs(os_.dt()) + 1e-300
        assert abs(hs.time() - os_.time()) <= 1e-15
    _compare(os_, hs)


def test_force_dt_and_min_dt(H):
    sc = _with(scenario(2, 4), "neumann")
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    _sync_state(os_, hs)
    assert abs(hs.min_dt() - os_.min_dt()) <= 1e-15 * os_.min_dt()
    os_.force_dt(1e-4)
    hs.dt(1e-4)
    os_.solve()
    hs.solve()
    assert hs.dt() == 1e-4 and os_.dt() == 1e-4
    _compare(os_, hs)


def test_conservation_sums_periodic(H):
    sc = _vortex()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    _sync_state(os_, hs)
    s0 = hs.conservation_sums()
    np.testing.assert_allclose(s0, os_.conservation_sums(), rtol=1e-13, atol=1e-13)
    hs.solve(5)
    s1 = hs.conservation_sums()
    # periodic: the telescoping face fluxes conserve every variable
    np.testing.assert_allclose(s1, s0, rtol=1e-12, atol=1e-12)


def test_check_state_flags_bad_cells(H):
    sc = _with(scenario(2, 4), "neumann")
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    q = os_.get("lhs").copy()
    assert hs.check_state() == 0
    q[2, 5] = -1.0       # negative density
    q[0, 7] = np.nan
    hs.set_lhs(q)
    assert hs.check_state() == 2


@pytest.mark.parametrize("name", ["cns3d_tgv_rk3_brick", "euler2d_vortex_periodic_muscl",
                                  "euler2d_neumann_rk3_brick"])
def test_brick_matches_generic(H, name):
    """The brick kernel (face-once, LDS) and the table-driven kernel agree."""
    sc = STEP_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hb = build_gpu(sc, kernel="brick")
    _, hg = build_gpu(sc, kernel="generic")
    assert hb.kernel() == "brick" and hg.kernel() == "generic"
    for hs in (hb, hg):
        _sync_state(os_, hs)
        hs.solve(5)
    n = os_.first_gc
    err = rel_l2(hb.lhs[:, :n], hg.lhs[:, :n], os_.nd)
    assert np.max(err) <= TOL, err


@pytest.mark.parametrize("name,tol", [("cns3d_tgv_rk3_brick", 1e-5), ("euler2d_vortex_periodic_muscl", 1e-5),
                                      ("heat2d_dirichlet", 1e-5), ("heat3d_neumann_rk2", 1e-5),
                                      ("cns2d_isothermal_walls_rk2", 1e-5), ("euler3d_dirichlet_rk3_8bricks", 1e-5)])
def test_fp32_state_error(H, name, tol):
    """FP32 stage state + FP64 RK accumulation against the FP64 oracle (the
    heat solver on the table-driven kernel, the others on bricks)."""
    sc = STEP_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc, precision="fp32")
    assert hs.kernel() == ("generic" if name.startswith("heat") else "brick")
    _sync_state(os_, hs)
    os_.solve(5)
    hs.solve(5)
    n = os_.first_gc
    qh, qo = hs.lhs[:, :n], os_.get("lhs")[:, :n]
    err = rel_l2(qh, qo, os_.nd)
    emax = np.max(np.abs(qh - qo), axis=1) / np.max(np.abs(qo), axis=1)
    print(name, "fp32 rel L2", err, "rel max", emax)
    assert np.max(err) <= tol, err


@pytest.mark.parametrize("name", ["cns3d_tgv_rk3_brick", "euler2d_vortex_periodic_muscl", "cns2d_isothermal_walls_rk2",
                                  "euler3d_dirichlet_rk3_8bricks"])
def test_fp32_state_table_driven(H, name):
    """The same scenarios on the table-driven kernel with FP32 stage states
    and FP32 fluxes: SSP-RK3 and Heun, MUSCL second neighbours across
    periodic images, wall slopes; relative L2 <= 1e-5 against the FP64
    oracle."""
    sc = STEP_SCENARIOS[name]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc, precision="fp32", kernel="generic")
    assert hs.kernel() == "generic"
    _sync_state(os_, hs)
    os_.solve(5)
    hs.solve(5)
    n = os_.first_gc
    qh, qo = hs.lhs[:, :n], os_.get("lhs")[:, :n]
    err = rel_l2(qh, qo, os_.nd)
    print(name, "table-driven fp32 rel L2", err)
    assert np.all(np.isfinite(qh))
    assert np.max(err) <= 1e-5, err
