"""Tangent and discrete adjoint of the conjugate heat transfer case against
central finite differences through the oracle (the CPU restatement of the
reference's coupled solver, coupling.hpp:47-63) on a frozen dt schedule.

Objective: time-integrated wall heat flux out of the solid; parameters: the
solid's per-cell volumetric source.  Checks: the GPU objective equals the
oracle's; FP64 tangent and adjoint = central FD to 1e-6 relative at eps
1e-4 through the oracle, and at eps 1e-6 through the oracle's long-double
build (in FP64 the difference of the two runs there carries a rounding
error of a few 1e-6 of dJ: 1e-5 is asserted for that one), along 16
seeded directions; adjoint dot
products = tangent to 1e-9 relative along 16 seeded directions; checkpoint
spacing does not change the gradient, nor does keeping the checkpoints in
pinned host memory, and two runs agree (bit for bit: the reverse sweep
sums in a fixed order); a finite end time that clamps the last steps is
handled by every checkpoint spacing; the torch.autograd.Function returns
the adjoint gradient.  The inverse-source objective (final-state misfit
against observed temperatures) gets the same checks and a few descent
steps of a source reconstruction."""
import numpy as np
import pytest

import oracle as O
from tests.test_gpu_cht import cht_setup, cht_setup_oracle

pytestmark = pytest.mark.gpu

TR0 = 0.01


def source_field(n):
    return np.random.default_rng(5).uniform(0, 1, n)


def setup(nd=2, level=5, ti="ssp_rk3", kratio=100.0):
    (oc, oh), (c, h) = cht_setup(nd, level, kratio=kratio, source=source_field, ti=ti)
    return oc, oh, c, h


def oracle_J(oc, oh, snap, src, nsteps, dts, Om=O):
    for s, q, t, dt, st in snap:
        s.set_lhs(q)
        s.set_time(t, dt, st)
    oh.set_source(src)
    J, _ = Om.cht_objective(TR0, oc, oh, nsteps, dts)
    return J


def oracle_snapshot(*solvers):
    return [(s, s.get("lhs").copy(), s.time(), s.dt(), s.step()) for s in solvers]


@pytest.mark.parametrize("ti,nsteps", [("ssp_rk3", 12), ("euler_forward", 10)])
def test_tangent_and_adjoint_match_finite_differences(ti, nsteps):
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup(ti=ti)
    osnap = oracle_snapshot(oc, oh)
    gsnap = A.Snapshot(c, h)
    n = oh.first_gc
    S0 = source_field(n)
    # reference schedule from the CFL-limited oracle run
    J_o, dts = O.cht_objective(TR0, oc, oh, nsteps)
    # GPU objective on the same schedule
    gsnap.restore()
    J_g, _ = A.cht_objective(c, h, TR0, nsteps, dts)
    assert abs(J_g - J_o) <= 1e-10 * abs(J_o), (J_g, J_o)
    # adjoint gradient
    gsnap.restore()
    J_a, grad, ms_f, ms_r = A.cht_adjoint(c, h, TR0, dts, ckpt_every=5)
    assert abs(J_a - J_o) <= 1e-10 * abs(J_o)
    assert np.all(np.isfinite(grad)) and np.any(grad != 0)
    rng = np.random.default_rng(123)
    # the oracle's long-double build from the same state on the same dt
    # schedule: the checker at eps = 1e-6
    OL = O.long_double()
    ocl, ohl = cht_setup_oracle(OL, nd=2, level=5, kratio=100.0, source=source_field, ti=ti)
    lsnap = [(sl,) + tuple(rest) for sl, (_, *rest) in zip((ocl, ohl), osnap)]
    J_l = oracle_J(ocl, ohl, lsnap, S0, nsteps, dts, OL)
    print(f"J fp64 oracle {J_o:.15e} long double {J_l:.15e}")
    assert abs(J_l - J_o) <= 1e-9 * abs(J_o), (J_l, J_o)

    def fd(v, eps):
        return (oracle_J(oc, oh, osnap, S0 + eps * v, nsteps, dts)
                - oracle_J(oc, oh, osnap, S0 - eps * v, nsteps, dts)) / (2 * eps)

    def fd_ld(v, eps):
        return (oracle_J(ocl, ohl, lsnap, S0 + eps * v, nsteps, dts, OL)
                - oracle_J(ocl, ohl, lsnap, S0 - eps * v, nsteps, dts, OL)) / (2 * eps)

    for k in range(16):
        v = rng.uniform(-1, 1, n)
        v /= np.linalg.norm(v)
        gsnap.restore()
        _, dJ_t = A.cht_tangent(c, h, TR0, dts, v)
        dJ_a = float(grad @ v)
        # eps = 1e-6: the difference of two FP64 objectives carries a
        # rounding error of about 1e-16 |J| / eps, a few 1e-6 of dJ here;
        # eps = 1e-4 leaves 100x less rounding and an O(eps^2) truncation
        fd6, fd4, fd6l = fd(v, 1e-6), fd(v, 1e-4), fd_ld(v, 1e-6)
        print(f"direction {k}: adjoint {dJ_a:.12e} tangent {dJ_t:.12e} fd(1e-6) {fd6:.12e} "
              f"fd(1e-4) {fd4:.12e} fd_ld(1e-6) {fd6l:.12e} rel {abs(dJ_a - fd6l) / abs(fd6l):.2e}")
        assert abs(dJ_a - dJ_t) <= 1e-9 * abs(dJ_t), (k, dJ_a, dJ_t)
        assert abs(dJ_t - fd4) <= 1e-6 * abs(fd4), (k, dJ_t, fd4)
        assert abs(dJ_a - fd4) <= 1e-6 * abs(fd4), (k, dJ_a, fd4)
        assert abs(dJ_a - fd6l) <= 1e-6 * abs(fd6l), (k, dJ_a, fd6l)
        assert abs(dJ_t - fd6l) <= 1e-6 * abs(fd6l), (k, dJ_t, fd6l)
        assert abs(dJ_a - fd6) <= 1e-5 * abs(fd6), (k, dJ_a, fd6)


def test_adjoint_dot_products_16_directions_and_checkpoint_spacing():
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup(nd=3, level=4)
    gsnap = A.Snapshot(c, h)
    n = h.first_gc
    nsteps = 7
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    grads = []
    for K in (1, 3, 7, 0):  # 0: every step stored, no recomputation
        gsnap.restore()
        _, g, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K)
        grads.append(g)
    # the recomputed segments repeat the forward steps bit for bit and the
    # reverse sweep sums in a fixed order: the same gradient for every spacing
    for g in grads[1:]:
        np.testing.assert_array_equal(g, grads[0])
    # checkpoints in pinned host memory: the same computation
    for K in (3, 1):
        gsnap.restore()
        _, gd, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K, ckpt_memory="device")
        gsnap.restore()
        _, gh, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K, ckpt_memory="host")
        np.testing.assert_array_equal(gh, gd, err_msg=str(K))
    rng = np.random.default_rng(2024)
    for k in range(16):
        v = rng.uniform(-1, 1, n)
        gsnap.restore()
        _, dJ_t = A.cht_tangent(c, h, TR0, dts, v)
        dJ_a = float(grads[0] @ v)
        assert abs(dJ_a - dJ_t) <= 1e-9 * abs(dJ_t), (k, dJ_a, dJ_t)


def test_torch_autograd_function():
    import torch
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup()
    gsnap = A.Snapshot(c, h)
    nsteps = 6
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    gsnap.restore()
    _, g_ref, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=3)
    S = torch.tensor(source_field(h.first_gc), dtype=torch.float64, requires_grad=True)
    Fn = A.WallHeatFlux()
    J = Fn.apply(S, c, h, TR0, dts, gsnap, 3)
    (2.0 * J).backward()
    np.testing.assert_allclose(S.grad.numpy(), 2.0 * g_ref, rtol=1e-12, atol=0)


def oracle_misfit(oc, oh, snap, src, nsteps, dts, obs):
    """Final-state misfit 1/2 sum (T_N - obs)^2 of the oracle's coupled run."""
    oracle_J(oc, oh, snap, src, nsteps, dts)
    T = oh.get("lhs")[0, :oh.first_gc]
    return 0.5 * float(np.sum((T - obs) ** 2))


def test_misfit_adjoint_matches_fd_and_tangent():
    """Inverse-source objective: misfit of the final solid temperature
    against observations made with another source; adjoint against the
    tangent (1e-9) and against central FD through the oracle (1e-6 at eps
    1e-4) along 8 seeded directions; storing every step and checkpointing
    give the same gradient."""
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup()
    osnap = oracle_snapshot(oc, oh)
    gsnap = A.Snapshot(c, h)
    n = oh.first_gc
    nsteps = 12
    S0 = source_field(n)
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    # observations: the final temperatures of the oracle run with a source
    # that differs from S0 by a smooth bump
    S_true = S0 + 0.5 * np.sin(np.linspace(0, 3 * np.pi, n))
    obs = np.empty(n)
    oracle_J(oc, oh, osnap, S_true, nsteps, dts)
    obs[:] = oh.get("lhs")[0, :n]
    J_o = oracle_misfit(oc, oh, osnap, S0, nsteps, dts, obs)
    gsnap.restore()
    h.set_source(S0)
    J_g, _ = A.cht_objective(c, h, TR0, nsteps, dts, obs=obs)
    assert J_o > 0 and abs(J_g - J_o) <= 1e-10 * J_o, (J_g, J_o)
    grads = []
    for K in (3, 0):
        gsnap.restore()
        h.set_source(S0)
        J_a, g, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K, obs=obs)
        assert abs(J_a - J_o) <= 1e-10 * J_o
        grads.append(g)
    np.testing.assert_array_equal(grads[1], grads[0])
    grad = grads[0]
    rng = np.random.default_rng(77)
    for k in range(8):
        v = rng.uniform(-1, 1, n)
        v /= np.linalg.norm(v)
        gsnap.restore()
        h.set_source(S0)
        _, dJ_t = A.cht_tangent(c, h, TR0, dts, v, obs=obs)
        dJ_a = float(grad @ v)
        eps = 1e-4
        fd = (oracle_misfit(oc, oh, osnap, S0 + eps * v, nsteps, dts, obs)
              - oracle_misfit(oc, oh, osnap, S0 - eps * v, nsteps, dts, obs)) / (2 * eps)
        print(f"misfit direction {k}: adjoint {dJ_a:.12e} tangent {dJ_t:.12e} fd(1e-4) {fd:.12e}")
        assert abs(dJ_a - dJ_t) <= 1e-9 * abs(dJ_t), (k, dJ_a, dJ_t)
        assert abs(dJ_a - fd) <= 1e-6 * abs(fd), (k, dJ_a, fd)


def test_source_reconstruction_descends():
    """A few steepest-descent steps on the source with the adjoint gradient
    (backtracking on the GPU objective) reduce the misfit against
    observations made with the true source; torch wrapper included."""
    import torch
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup()
    gsnap = A.Snapshot(c, h)
    n = h.first_gc
    nsteps = 10
    S0 = source_field(n)
    gsnap.restore()
    _, dts = A.cht_objective(c, h, TR0, nsteps)
    S_true = S0 + 0.5 * np.cos(np.linspace(0, 2 * np.pi, n))
    gsnap.restore()
    h.set_source(S_true)
    A.cht_objective(c, h, TR0, nsteps, dts)
    obs = h.lhs[0, :n].copy()
    Fn = A.SourceMisfit()

    def misfit(S):
        gsnap.restore()
        h.set_source(S)
        return A.cht_objective(c, h, TR0, nsteps, dts, obs=obs)[0]

    S = torch.tensor(S0, dtype=torch.float64, requires_grad=True)
    J0 = None
    for it in range(3):
        J = Fn.apply(S, c, h, TR0, dts, obs, gsnap, 0)
        J.backward()
        g = S.grad.detach().numpy().copy()
        S.grad = None
        if J0 is None:
            J0 = float(J)
        # backtracking along -g
        alpha = float(J) / max(float(g @ g), 1e-300)
        Jc = float(J)
        for _ in range(30):
            Jn = misfit(S.detach().numpy() - alpha * g)
            if Jn < Jc - 1e-4 * alpha * float(g @ g):
                break
            alpha *= 0.5
        assert Jn < Jc, (it, Jn, Jc)
        with torch.no_grad():
            S -= alpha * torch.from_numpy(g)
        print(f"iteration {it}: misfit {Jc:.6e} -> {Jn:.6e} (step {alpha:.3e})")
    assert Jn < 0.9 * J0, (Jn, J0)


def test_checkpoint_memory_setting_rejects_unknown_placement():
    import hom3_amd as H
    from hom3_amd import _lib as L
    ctx = H.default_context()
    lib = L.lib()
    for where in (0, 1, 2):
        assert lib.hom3g_set_checkpoint_memory(ctx.h, where) == 0
    assert lib.hom3g_set_checkpoint_memory(ctx.h, 3) != 0
    assert b"checkpoint memory" in lib.hom3g_last_error()
    assert lib.hom3g_set_checkpoint_memory(ctx.h, 0) == 0


@pytest.mark.parametrize("nd,level", [(2, 5), (3, 4)])
def test_adjoint_is_bit_reproducible(nd, level):
    """Two runs of the same adjoint give the same gradient and objective bit
    for bit (no atomics in the reverse sweep or the objective sums)."""
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup(nd=nd, level=level)
    gsnap = A.Snapshot(c, h)
    nsteps = 6
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    out = []
    for K in (2, 2, 0, 0):
        gsnap.restore()
        J, g, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K)
        out.append((J, g))
    for J, g in out[1:]:
        assert J == out[0][0]
        np.testing.assert_array_equal(g, out[0][1])
    gsnap.restore()
    J1, dJ1 = A.cht_tangent(c, h, TR0, dts, np.ones(h.first_gc))
    gsnap.restore()
    J2, dJ2 = A.cht_tangent(c, h, TR0, dts, np.ones(h.first_gc))
    assert J1 == J2 == out[0][0] and dJ1 == dJ2


def test_adjoint_with_a_clamping_end_time():
    """timeEnd reached inside the run: set_dt clamps the step that crosses it
    (solver.hpp:541-551) and later steps take dt = 0.  The checkpointed
    recomputation restores each checkpoint's time and step with its state, and
    the reverse sweep uses the dts the solvers took: every spacing gives the
    stored-steps gradient, which matches central differences through the
    oracle run with the same end time."""
    from hom3_amd import adjoint as A
    nsteps = 12
    oc, oh, c, h = setup()
    _, dts = O.cht_objective(TR0, oc, oh, nsteps)
    t_end = float(np.sum(dts[:7]) + 0.5 * dts[7])
    scale = TR0 * 0.72 * oc.quantities()["Re0"]
    (oc, oh), (c, h) = cht_setup(2, 5, kratio=100.0, source=source_field, ti="ssp_rk3",
                                 time_end=(t_end, t_end * scale))
    osnap = oracle_snapshot(oc, oh)
    gsnap = A.Snapshot(c, h)
    J_o, used = O.cht_objective(TR0, oc, oh, nsteps, dts)
    assert used[7] < dts[7] and np.all(used[8:] == 0.0)
    grads = []
    for K in (0, 5, 3):
        gsnap.restore()
        J, g, _, _ = A.cht_adjoint(c, h, TR0, dts, ckpt_every=K)
        assert abs(J - J_o) <= 1e-10 * abs(J_o), (K, J, J_o)
        assert abs(c.time() - t_end) <= 1e-14 * t_end
        grads.append(g)
    for g in grads[1:]:
        np.testing.assert_array_equal(g, grads[0])
    n = oh.first_gc
    S0 = source_field(n)
    rng = np.random.default_rng(9)
    for k in range(4):
        v = rng.uniform(-1, 1, n)
        fd = (oracle_J(oc, oh, osnap, S0 + 1e-4 * v, nsteps, dts)
              - oracle_J(oc, oh, osnap, S0 - 1e-4 * v, nsteps, dts)) / 2e-4
        dJ = float(grads[0] @ v)
        assert abs(dJ - fd) <= 1e-6 * abs(fd), (k, dJ, fd)
        gsnap.restore()
        _, dJ_t = A.cht_tangent(c, h, TR0, dts, v)
        assert abs(dJ - dJ_t) <= 1e-9 * abs(dJ_t), (k, dJ, dJ_t)


def test_restore_then_solve_matches_a_fresh_run():
    """reset_time drops a dt forced before it: after an adjoint run of 4
    steps (which forces the dt of steps 0-3) and a reset to step 3, an
    uncoupled step takes the CFL dt, as a freshly set-up solver does."""
    from hom3_amd import adjoint as A
    oc, oh, c, h = setup()
    gsnap = A.Snapshot(c, h)
    _, dts = O.cht_objective(TR0, oc, oh, 4)
    gsnap.restore()
    A.cht_adjoint(c, h, TR0, dts, ckpt_every=2)
    _, _, c2, h2 = setup()
    for s, snap in ((c, gsnap.items[0]), (c2, gsnap.items[0])):
        s.set_lhs(snap[1])
        s.reset_time(snap[2], snap[3], 3)
    h.set_lhs(gsnap.items[1][1])
    c.solve(1)
    c2.solve(1)
    assert c.dt() == c2.dt() and c.step() == c2.step() == 4
    np.testing.assert_array_equal(c.lhs, c2.lhs)
