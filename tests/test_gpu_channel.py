"""The channel-over-a-heated-block case (fluid channel, solid block in the
lower wall, conjugate top face) at a reduced level: GPU topology and coupled
steps against the oracle, and the adjoint of the wall heat flux against
central finite differences along 16 seeded directions: through the oracle
at eps 1e-4, and at eps 1e-6 through the oracle's long-double build (to
1e-6 relative both)."""
import numpy as np
import pytest

import oracle as O
from tests.scenarios import rel_l2

pytestmark = pytest.mark.gpu

TR0 = 0.01


def oracle_channel(level, kratio=100.0, seed=7, ti=O.TI_SSP_RK3, Minfty=0.2, ReInfty=1000.0, Pr0=0.71, cfl=0.4,
                   heat_cfl=0.3, O=O):
    from hom3_amd.cases import CHANNEL as C
    nd = 3
    h = 1.0 / 2 ** level
    g = O.Grid(nd, [0.0] * 3, [1.0] * 3, max_solvers=2).min_level(level)
    cns = O.Solver(g, O.phys(physics=O.PHYS_CNS, nd=nd, ti=ti, cfl=cfl, cfl_viscous=cfl, Minfty=Minfty,
                             ReInfty=ReInfty, Pr0=Pr0), 0)
    heat = O.Solver(g, O.phys(physics=O.PHYS_HEAT, nd=nd, flux=O.FLUX_THREE_POINT, ti=ti, cfl=heat_cfl), 1)
    q = cns.quantities()
    pv = [q["u_inf"], 0.0, 0.0, q["rho_inf"], q["p_inf"]]
    fluid = O.geom(O.GEOM_SQUARE, C["fluid_centre"], C["fluid_dims"], invert=True)
    block = O.geom(O.GEOM_SQUARE, C["block_centre"], C["block_dims"], invert=True)
    cube1 = O.geom(O.GEOM_SQUARE, C["block_centre"], [d + h for d in C["block_dims"]])
    cube2 = O.geom(O.GEOM_SQUARE, C["block_centre"], [d - 0.6 * h for d in C["block_dims"]], invert=True)
    cube = O.cube_boundaries(nd, [0.0] * 3, [1.0] * 3)
    cns.append_bc(cube[0], O.BC_CNS_INFLOW, pv)
    cns.append_bc(cube[1], O.BC_CNS_OUTFLOW, pv)
    cns.append_bc(cube1, O.BC_CHT_FLUID, [], other=heat, kratio=kratio)
    cns.append_bc(O.geom(O.GEOM_EDGE, [0.5, 0.25, 0.25], [0.0, 1.0, 0.0]), O.BC_ADIABATIC_NOSLIP, [])
    cns.append_bc(O.geom(O.GEOM_EDGE, [0.5, 0.75, 0.25], [0.0, -1.0, 0.0]), O.BC_ADIABATIC_NOSLIP, [])
    cns.append_bc(cube[4], O.BC_ADIABATIC_NOSLIP, [])
    cns.append_bc(O.geom(O.GEOM_EDGE, [0.5, 0.5, 0.5], [0.0, 0.0, -1.0]), O.BC_ADIABATIC_NOSLIP, [])
    heat.append_bc(O.geom(O.GEOM_EDGE, [0.5, 0.25, 0.25], [0.0, -1.0, 0.0]), O.BC_CHT_SOLID, [], other=cns,
                   kratio=kratio)
    heat.append_bc(cube2, O.BC_HEAT_NEUMANN, [0.0])
    cns.set_domain(fluid, +1)
    heat.set_domain(block, +1)
    cns.set_ic(O.IC_CONSTANT_PV, pv)
    heat.set_ic(O.IC_HEAT_CONSTANT, [q["T_inf"]])
    heat.initialize()
    cns.initialize()
    src = np.random.default_rng(seed).uniform(0.0, 1.0, heat.first_gc)
    heat.set_source(src)
    return g, cns, heat, src


def build(level):
    import hom3_amd as H
    from hom3_amd.cases import channel_cht
    og, oc, oh, osrc = oracle_channel(level)
    g, c, h, src = channel_cht(H, level=level)
    np.testing.assert_array_equal(src, osrc)
    for o, s in ((oc, c), (oh, h)):
        s.set_lhs(o.get("lhs"))
        s.reset_time(o.time(), o.dt(), o.step())
    return (oc, oh), (c, h)


def test_channel_topology_and_steps():
    (oc, oh), (c, h) = build(4)
    for o, s in ((oc, c), (oh, h)):
        t = s.topology()
        assert s.ncells == o.ncells and s.first_gc == o.first_gc
        for key in ("neighbors", "node_idx", "bc_idx", "x_center", "length"):
            np.testing.assert_array_equal(t[key], o.get(key), err_msg=key)
    # 16 x 8 x 8 fluid and 4 x 2 x 4 solid cells at level 4
    assert oc.first_gc == 16 * 8 * 8 and oh.first_gc == 4 * 2 * 4
    import hom3_amd as H
    nsteps = 5
    for _ in range(nsteps):
        O.couple_cns_heat(TR0, oc, oh)
        oc.solve()
        oh.solve()
    H.coupling.apply_and_step(TR0, c, h, nsteps)
    for o, s in ((oc, c), (oh, h)):
        n = o.first_gc
        err = rel_l2(s.lhs[:, :n], o.get("lhs")[:, :n], 3 if o.nvars > 1 else None)
        assert np.max(err) <= 1e-12, err


def test_channel_adjoint_matches_finite_differences():
    from hom3_amd import adjoint as A
    (oc, oh), (c, h) = build(4)
    snap = [(s, s.get("lhs").copy(), s.time(), s.dt(), s.step()) for s in (oc, oh)]
    gsnap = A.Snapshot(c, h)
    S0 = np.random.default_rng(7).uniform(0.0, 1.0, oh.first_gc)
    nsteps = 8
    J_o, dts = O.cht_objective(TR0, oc, oh, nsteps)
    gsnap.restore()
    J_a, grad, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=3)
    assert abs(J_a - J_o) <= 1e-10 * abs(J_o)

    # the long-double oracle from the same state, on the same dt schedule
    OL = O.long_double()
    _, ocl, ohl, _ = oracle_channel(4, O=OL)
    lsnap = [(sl,) + tuple(rest) for sl, (_, *rest) in zip((ocl, ohl), snap)]

    def oracle_J(src, Om=O, sn=snap):
        a, b = sn[0][0], sn[1][0]
        for s, q, t, dt, st in sn:
            s.set_lhs(q)
            s.set_time(t, dt, st)
        b.set_source(src)
        return Om.cht_objective(TR0, a, b, nsteps, dts)[0]

    assert abs(oracle_J(S0, OL, lsnap) - J_o) <= 1e-9 * abs(J_o)
    rng = np.random.default_rng(16)
    for k in range(16):
        v = rng.uniform(-1, 1, S0.size)
        dJ = grad @ v
        fd4 = (oracle_J(S0 + 1e-4 * v) - oracle_J(S0 - 1e-4 * v)) / 2e-4
        fd6 = (oracle_J(S0 + 1e-6 * v, OL, lsnap) - oracle_J(S0 - 1e-6 * v, OL, lsnap)) / 2e-6
        print(f"direction {k}: adjoint {dJ:.12e} fd_ld(1e-6) {fd6:.12e} fd(1e-4) {fd4:.12e} "
              f"rel {abs(dJ - fd6) / abs(fd6):.2e}")
        assert abs(dJ - fd4) <= 1e-6 * abs(fd4), (k, dJ, fd4)
        assert abs(dJ - fd6) <= 1e-6 * abs(fd6), (k, dJ, fd6)


def test_channel_checkpoints_every_20_steps():
    """The configured checkpoint spacing (every 20 steps) over 60 coupled
    steps: the gradient equals the one with every step stored, and agrees
    with central differences through the oracle along 4 seeded directions."""
    from hom3_amd import adjoint as A
    (oc, oh), (c, h) = build(4)
    snap = [(s, s.get("lhs").copy(), s.time(), s.dt(), s.step()) for s in (oc, oh)]
    gsnap = A.Snapshot(c, h)
    S0 = np.random.default_rng(7).uniform(0.0, 1.0, oh.first_gc)
    nsteps = 60
    J_o, dts = O.cht_objective(TR0, oc, oh, nsteps)
    gsnap.restore()
    J20, g20, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=20)
    assert c.step() == nsteps and h.step() == nsteps
    gsnap.restore()
    J0, g0, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=0)
    assert J20 == J0 and abs(J20 - J_o) <= 1e-10 * abs(J_o)
    np.testing.assert_array_equal(g20, g0)

    def oracle_J(src):
        for s, q, t, dt, st in snap:
            s.set_lhs(q)
            s.set_time(t, dt, st)
        oh.set_source(src)
        return O.cht_objective(TR0, oc, oh, nsteps, dts)[0]
    rng = np.random.default_rng(20)
    for k in range(4):
        v = rng.uniform(-1, 1, S0.size)
        fd = (oracle_J(S0 + 1e-4 * v) - oracle_J(S0 - 1e-4 * v)) / 2e-4
        dJ = float(g20 @ v)
        print(f"direction {k}: adjoint {dJ:.12e} fd(1e-4) {fd:.12e} rel {abs(dJ - fd) / abs(fd):.2e}")
        assert abs(dJ - fd) <= 1e-6 * abs(fd), (k, dJ, fd)


def test_channel_binomial_checkpoints():
    """Binomial checkpointing (ckpt_every = -S: at most S saved states) over
    24 coupled steps: the gradient and J equal those with every step stored,
    bit for bit, for S = 1, 3 and 8, for the wall heat flux and for the
    inverse-source misfit; the solvers end at step 24."""
    from hom3_amd import adjoint as A
    (oc, oh), (c, h) = build(4)
    gsnap = A.Snapshot(c, h)
    nsteps = 24
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    T_obs = np.random.default_rng(3).uniform(0.9, 1.1, h.first_gc) * float(np.mean(h.lhs[0, :h.first_gc]))
    for obs in (None, T_obs):
        gsnap.restore()
        J0, g0, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=0, obs=obs)
        for S in (1, 3, 8):
            gsnap.restore()
            JS, gS, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=-S, obs=obs)
            assert JS == J0, (S, JS, J0)
            np.testing.assert_array_equal(gS, g0)
            assert c.step() == nsteps and h.step() == nsteps
