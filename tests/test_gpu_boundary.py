"""The drop-in boundary on the GPU: initial conditions and boundary
conditions given as host code, the C++ facade and the reference-side
binding driving the library, VTK output of device state, apply_bcs and the
divergence report of solve(); each against the oracle.

* a Python callable initial condition (solver.hpp:153-156) is averaged at
  the Gauss points by the library exactly as the oracle averages the same
  callable (bit for bit);
* host-evaluated Conditions (HOM3G_BC_HOST_CALLBACK): a Dirichlet condition
  written in Python against the oracle's Dirichlet, a wall condition with the
  wall slope hook against the oracle's adiabatic no-slip wall, an
  IsentropicVortex written in Python (it overwrites boundary cells) against
  the oracle's vortex condition;
* tests/cpp/vortex_driver transcribes euler_fv_solver_test.cpp:258-286 (a
  lambda initial condition, the IsentropicVortex condition,
  run_solver(grid, solver, maxNoTimeSteps, outputInterval)) against the
  facade, with the built-in condition and with the same condition written in
  C++ on the host; its VTK files hold the device state;
* tests/cpp/binding_check builds the solver from an io::Properties map
  through the reference-side binding (tests/cpp/hom3_binding.hpp).
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests.scenarios import build_gpu, build_oracle, rel_l2, scenario

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def _sync(os_, hs):
    hs.set_lhs(os_.get("lhs"))
    hs.reset_time(os_.time(), os_.dt(), os_.step())


def _close(os_, hs, tol=1e-12):
    n = os_.first_gc
    err = rel_l2(hs.lhs[:, :n], os_.get("lhs")[:, :n], os_.nd if os_.nvars > 1 else None)
    assert np.all(np.isfinite(hs.lhs[:, :n]))
    assert np.max(err) <= tol, err
    return err


def test_callable_initial_condition_matches_oracle_bit_for_bit():
    import hom3_amd as H

    def f(x):  # a smooth 3D state, conserved variables
        rho = 1.0 + 0.2 * math.sin(2 * x[0]) * math.cos(x[1]) + 0.1 * x[2]
        u, v, w = 0.1 * math.cos(x[1]), -0.05 * x[0], 0.02
        p = 1.0 + 0.1 * math.cos(x[2])
        return [rho * u, rho * v, rho * w, rho, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)]
    sc = scenario(3, 3, CFL=0.3)
    sc["bcs"] = [("cube", i, O.BC_EULER_NEUMANN, []) for i in range(6)]
    sc["ic"] = (O.IC_CALLBACK, [], f)
    _, os_ = build_oracle(sc)
    g = H.Grid(3, [0.0] * 3, [1.0] * 3, 3)
    s = H.Solver(g, physics="euler", CFL=0.3)
    H.append_bcs(s, H.euler.bc.Neumann())
    s.set_initial_condition(f)
    s.initialize()
    n = os_.first_gc
    np.testing.assert_array_equal(s.lhs[:, :n], os_.get("lhs")[:, :n])
    # the CFL kernels use refined reciprocals: the minimum agrees to rounding
    assert abs(s.dt() - os_.dt()) <= 1e-14 * os_.dt()


def _host_dirichlet(values):
    vals = np.asarray(values, float)

    def fill(gs):  # boundary_condition.hpp:70-72: ghost = 2 s - b
        gs.q_ghost[:] = 2.0 * vals[:, None] - gs.q_bndry
    return fill


@pytest.mark.parametrize("ti", ["euler_forward", "ssp_rk3"])
def test_host_dirichlet_condition(ti):
    import hom3_amd as H
    nd = 2
    cv = [0.1, 0.05, 1.0, 2.5 + 0.5 * 1.0 * (0.01 + 0.0025)]
    sc = scenario(nd, 5, ti=ti, CFL=0.3)
    sc["bcs"] = [("cube", i, O.BC_EULER_DIRICHLET, cv) for i in range(2 * nd)]
    sc["ic"] = (O.IC_SPHERE, [0.5, 0.5, 0.25, 0.1, 0.05, 1.3, 1.5, 0.1, 0.05, 1.0, 1.0])
    _, os_ = build_oracle(sc)
    g = H.Grid(nd, [0.0] * nd, [1.0] * nd, 5)
    s = H.Solver(g, physics="euler", time_integration=ti, CFL=0.3)
    H.append_bcs(s, H.HostCondition(_host_dirichlet(cv)))
    s.set_initial_condition(O.IC_SPHERE, sc["ic"][1])
    s.initialize()
    assert s.kernel() == "generic"
    _sync(os_, s)
    os_.solve(6)
    s.solve(6)
    _close(os_, s)


def test_host_wall_condition_with_wall_slope_hook():
    """A no-slip adiabatic wall written on the host (u_g = -u_b, rho and p
    mirrored) with the wall slope hook equals the built-in wall on a
    Navier-Stokes run (the hook enters the viscous tangential slopes)."""
    import hom3_amd as H
    from tests.test_gpu_parity import _cns_sod
    sc = _cns_sod(2, 5, "wall", "runge_kutta_2")
    _, os_ = build_oracle(sc)
    nd = 2

    def wall(gs):
        rho = gs.q_bndry[nd]
        gs.q_ghost[:nd] = -gs.q_bndry[:nd]          # rho u_g = -rho u_b at equal rho
        gs.q_ghost[nd] = rho
        gs.q_ghost[nd + 1] = gs.q_bndry[nd + 1]     # same p and |u|: same rho E
    g = H.Grid(nd, [0.0] * nd, [1.0] * nd, 5)
    p = dict(sc["props"])
    s = H.Solver(g, physics="cns", time_integration="runge_kutta_2", **p)
    H.append_bcs(s, H.HostCondition(wall, hook="wall"))
    s.set_initial_condition(sc["ic"][0], sc["ic"][1])
    s.initialize()
    _sync(os_, s)
    os_.solve(6)
    s.solve(6)
    _close(os_, s)
    # the default hook gives a different (wrong) viscous flux at the wall
    s2 = H.Solver(H.Grid(nd, [0.0] * nd, [1.0] * nd, 5), physics="cns", time_integration="runge_kutta_2", **p)
    H.append_bcs(s2, H.HostCondition(wall, hook="default"))
    s2.set_initial_condition(sc["ic"][0], sc["ic"][1])
    s2.initialize()
    _, o2 = build_oracle(sc)
    _sync(o2, s2)
    s2.solve(6)
    n = o2.first_gc
    assert np.max(rel_l2(s2.lhs[:, :n], os_.get("lhs")[:, :n], nd)) > 1e-10


def _vortex_oracle(level=5, ti=O.TI_HEUN_RK2):
    og = O.Grid(2, [-8.0, -8.0], [8.0, 8.0]).min_level(level)
    os_ = O.Solver(og, O.phys(nd=2, ti=ti, cfl=0.2, time_end=2.0))
    os_.append_cube_bcs([O.BC_ISENTROPIC_VORTEX] * 4)
    os_.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0])
    os_.initialize()
    return og, os_


def test_host_vortex_condition_overwrites_boundary_cells():
    import hom3_amd as H
    og, os_ = _vortex_oracle()
    g = H.Grid(2, [-8.0, -8.0], [8.0, 8.0], 5)
    s = H.Solver(g, physics="euler", time_integration="runge_kutta_2", CFL=0.2, timeEnd=2.0)

    def vortex(gs):  # the analytic vortex at the ghost and boundary centres
        for q, x in ((gs.q_ghost, gs.x_ghost), (gs.q_bndry, gs.x_bndry)):
            for i in range(gs.n):
                q[:, i] = os_.ic_point(x[:, i] - gs.time)
    H.append_bcs(s, H.HostCondition(vortex))
    s.set_initial_condition(O.IC_ISENTROPIC_VORTEX, [0.0])
    s.initialize()
    _sync(os_, s)
    os_.solve(5)
    s.solve(5)
    _close(os_, s)


def _run(exe, *args, cwd=None):
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    r = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True, timeout=300, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.split("\n")
    head = lines[0].split()
    n, nint, nv = int(head[0]), int(head[1]), int(head[2])
    q = np.array([float(x) for x in lines[1:1 + nv * nint]]).reshape(nv, nint)
    return head, q


@pytest.mark.parametrize("bc", ["builtin", "host"])
def test_vortex_driver_through_the_facade(bc, tmp_path):
    """euler_fv_solver_test.cpp:258-286 against the facade: 20 steps of the
    vortex (lambda IC, IsentropicVortex condition, forward Euler, CFL 0.2,
    level 5 on [-8, 8]^2) match the oracle; with an output interval of 10 the
    VTK files of steps 0, 10 and 20 hold the device state."""
    from hom3_amd import vtk
    steps = 20
    head, q = _run("vortex_driver", steps, 10, *(["host"] if bc == "host" else []), cwd=str(tmp_path))
    og, os_ = _vortex_oracle(ti=O.TI_EULER_FWD)
    os_.solve(steps)
    # run_solver ends with write_domain, which applies the conditions to lhs
    # first (solver.hpp:329-333): the vortex condition rewrites the boundary
    # cells with the analytic state at the final time
    os_.apply_bcs_lhs()
    n = os_.first_gc
    assert int(head[1]) == n and int(head[5]) == steps
    err = rel_l2(q, os_.get("lhs")[:, :n], 2)
    assert np.max(err) <= 1e-12, err
    assert abs(float(head[3]) - os_.time()) <= 1e-13 * os_.time()
    files = sorted(os.listdir(tmp_path))
    assert files == ["Fv_Euler2D_Id0_0.vtk", "Fv_Euler2D_Id0_10.vtk", "Fv_Euler2D_Id0_20.vtk"], files
    qo = os_.get("lhs")
    _, cells, data = vtk.read_cell_data(str(tmp_path / "Fv_Euler2D_Id0_20.vtk"))
    for v, name in enumerate(["rho_u0", "rho_u1", "rho", "rho_E"]):
        np.testing.assert_allclose(data[name], qo[v].astype(np.float32), rtol=2.0 ** -23, atol=0)
    np.testing.assert_array_equal(data["gridNodeIds"][:n], os_.get("node_idx")[0][:n])
    assert np.all(data["ghost_cell"][n:] == 1) and np.all(data["ghost_cell"][:n] == -1)


def test_binding_from_properties_map():
    head, q = _run("binding_check", 4)
    g = O.Grid(3, [0.0] * 3, [1.0] * 3).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_EULER, nd=3, cfl=0.2, time_end=0.2))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    s.set_ic(O.IC_SHOCK_TUBE, [0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1])
    s.initialize()
    s.solve(4)
    err = rel_l2(q, s.get("lhs")[:, :s.first_gc], 3)
    assert np.max(err) <= 1e-12, err


def test_vtk_of_device_state_after_steps(tmp_path):
    """hom3_amd.vtk writes the GPU solver's state (io/vtk.hpp:363-364:
    float32 values): the cell arrays equal the oracle state cast to float32."""
    from hom3_amd import vtk
    from tests.test_gpu_parity import STEP_SCENARIOS
    sc = STEP_SCENARIOS["cns3d_sod_slipwalls_rk3"]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    _sync(os_, hs)
    os_.solve(4)
    hs.solve(4)
    hs.apply_bcs()
    os_.apply_bcs_lhs()
    path = vtk.write_domain(str(tmp_path / "cns"), hs)
    _, cells, data = vtk.read_cell_data(path)
    qo = os_.get("lhs")
    # transverse momenta are round-off (1e-22) in this tube: their scale is
    # the momentum field's
    amom = 1e-12 * np.max(np.abs(qo[:3]))
    for v, name in enumerate(["rho_u0", "rho_u1", "rho_u2", "rho", "rho_E"]):
        np.testing.assert_allclose(data[name], qo[v].astype(np.float32), rtol=2.0 ** -23,
                                   atol=amom if v < 3 else 0)
    nb = os_.get("neighbors")
    for p in range(6):
        ref = np.where(nb[p] == O.INVALID, -1, nb[p])
        np.testing.assert_array_equal(data[f"local_nghbrs_{p}"], ref)


def test_apply_bcs_matches_oracle():
    from tests.test_gpu_parity import STEP_SCENARIOS
    sc = STEP_SCENARIOS["cns2d_isothermal_walls_rk2"]()
    _, os_ = build_oracle(sc)
    _, hs = build_gpu(sc)
    q = os_.get("lhs")
    n = os_.first_gc
    q[:, :n] *= 1 + 0.01 * np.random.default_rng(1).standard_normal(q[:, :n].shape)
    os_.set_lhs(q)
    hs.set_lhs(q)
    os_.apply_bcs_lhs()
    hs.apply_bcs()
    np.testing.assert_allclose(hs.lhs[:, n:], os_.get("lhs")[:, n:], rtol=1e-14, atol=1e-15)


def test_solve_reports_divergence():
    """utilities.hpp:73-76, 107-110: a dt outside [1e-300, 1e300] is
    "Solution diverged!" (HOM3G_ERR_DIVERGED from solve, SolutionDiverged in
    Python); reset_time clears the report."""
    import hom3_amd as H
    from tests.test_gpu_parity import STEP_SCENARIOS
    sc = STEP_SCENARIOS["euler2d_sod_fe"]()
    _, hs = build_gpu(sc)
    hs.solve(2)
    hs.dt(1e-301)
    with pytest.raises(H.SolutionDiverged) as e:
        hs.solve(3)
    assert "step 2" in str(e.value)
    hs.reset_time(0.0, 0.0, 0)
    hs.solve(2)
    assert hs.step() == 2
