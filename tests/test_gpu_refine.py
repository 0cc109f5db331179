"""2:1 refined grids on the GPU against the oracle: tree (refinement and
balance passes), cells, neighbours with coarse and fine faces, ghosts,
seeded perturbation and first-order Euler AUSM steps."""
import numpy as np
import pytest

import oracle as O
from tests.scenarios import rel_l2

pytestmark = pytest.mark.gpu

# (nd, base, extra, centre, radius, band, ic, nsteps, ti, cfl)
BLAST3 = [0.5, 0.5, 0.5, 0.15, 0.0, 0.0, 0.0, 1.0, 10.0, 0.0, 0.0, 0.0, 0.125, 0.1]
QUAD = [0.4, 0.6, 0.1, 0.2, 1.0, 1.0, -0.3, 0.1, 0.5, 0.4, 0.0, 0.0, 2.0, 1.5, 0.2, -0.2, 0.8, 0.9]
CASES = {
    "blast3d_l3+2": (3, 3, 2, [0.5, 0.5, 0.5], 0.15, 0.025, (O.IC_SPHERE, BLAST3), 4, "ssp_rk3", 0.1),
    "quad2d_l5+2": (2, 5, 2, [0.35, 0.6], 0.2, 0.01, (O.IC_FOUR_QUADRANT, QUAD), 5, "ssp_rk3", 0.4),
    "quad2d_l4+1_fe": (2, 4, 1, [0.7, 0.3], 0.25, 0.01, (O.IC_FOUR_QUADRANT, QUAD), 6, "euler_forward", 0.4),
}
TI = {"euler_forward": O.TI_EULER_FWD, "ssp_rk3": O.TI_SSP_RK3}


def build_pair(name, seed=7, amp=0.01, precision="fp64"):
    import hom3_amd as H
    nd, base, extra, centre, radius, band, ic, nsteps, ti, cfl = CASES[name]
    xmin, xmax = [0.0] * nd, [1.0] * nd
    og = O.Grid(nd, xmin, xmax).min_level(base)
    og.refine_region(base + extra, O.geom(O.GEOM_SPHERE, centre, [radius]), band)
    og.balance()
    os_ = O.Solver(og, O.phys(physics=O.PHYS_EULER, nd=nd, ti=TI[ti], cfl=cfl))
    os_.append_cube_bcs([O.BC_EULER_NEUMANN] * (2 * nd))
    os_.set_ic(*ic)
    os_.set_perturbation(seed, amp)
    os_.initialize()

    g = H.Grid(nd, xmin, xmax, base)
    nref = g.refine_region(base + extra, H.Sphere(centre, radius), band)
    nbal = g.balance()
    s = H.Solver(g, physics="euler", time_integration=ti, CFL=cfl, precision=precision)
    H.append_bcs(s, H.euler.bc.Neumann())
    s.set_initial_condition(*ic)
    s.set_perturbation(seed, amp)
    s.initialize()
    return (og, os_), (g, s), nsteps, (nref, nbal)


@pytest.mark.parametrize("name", list(CASES))
def test_refined_topology_bit_exact(name):
    (og, o), (g, s), _, (nref, nbal) = build_pair(name)
    assert nref > 0
    po, fo = og.tree()
    pg, fg = g.tree()
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(fg, fo)
    assert s.ncells == o.ncells and s.first_gc == o.first_gc
    t = s.topology()
    for key in ("neighbors", "node_idx", "bc_idx", "x_center", "length"):
        np.testing.assert_array_equal(t[key], o.get(key), err_msg=key)
    fo_ = o.fine()
    fg_ = s.fine_faces()
    assert fo_.shape[0] > 0
    np.testing.assert_array_equal(fg_.astype(np.int64), fo_)
    assert s.kernel() == "generic"


@pytest.mark.parametrize("name", list(CASES))
def test_refined_steps_match_oracle(name):
    (og, o), (g, s), nsteps, _ = build_pair(name)
    nd = o.nd
    n = o.first_gc
    # initial condition + perturbation
    err0 = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], nd)
    assert np.max(err0) <= 1e-14, err0
    o.solve(nsteps)
    s.solve(nsteps)
    err = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], nd)
    assert np.max(err) <= 1e-12, err
    assert abs(s.time() - o.time()) <= 1e-13 * o.time()
    # conservation of the volume sums is the oracle's; GPU sums agree
    np.testing.assert_allclose(s.conservation_sums(), o.conservation_sums(), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("name", list(CASES))
def test_refined_fp32_state_against_oracle(name):
    """FP32 stage states on the table-driven kernel of a refined grid (the
    stage buffers and the fluxes in FP32, the RK combination and U^n in
    FP64):
    relative L2 <= 1e-5 per variable against the FP64 oracle after the
    case's steps (forward Euler has no stage buffers: the FP64 result)."""
    (og, o), (g, s), nsteps, _ = build_pair(name, precision="fp32")
    assert s.kernel() == "generic"
    nd = o.nd
    n = o.first_gc
    o.solve(nsteps)
    s.solve(nsteps)
    err = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], nd)
    print("FP32 refined", name, "rel L2", err)
    assert np.all(np.isfinite(s.lhs[:, :n]))
    assert np.max(err) <= 1e-5, err
    assert abs(s.time() - o.time()) <= 1e-6 * o.time()
