"""Tangent and discrete adjoint of the conjugate heat transfer case against
central finite differences through the oracle (the CPU restatement of the
reference's coupled solver, coupling.hpp:47-63) on a frozen dt schedule.

Objective: time-integrated wall heat flux out of the solid; parameters: the
solid's per-cell volumetric source.  Checks: the GPU objective equals the
oracle's; FP64 tangent and adjoint = central FD to 1e-6 relative at eps
1e-4 and to 1e-5 at eps 1e-6 (where the FD's own rounding, ~1e-16 |J| /
eps, is a few 1e-6 of dJ); adjoint dot products = tangent to 1e-9 relative
along 16 seeded directions;
checkpoint spacing does not change the gradient; the torch.autograd.Function
returns the adjoint gradient."""
import numpy as np
import pytest

import oracle as O
from tests.test_gpu_cht import cht_setup

pytestmark = pytest.mark.gpu

TR0 = 0.01


def source_field(n):
    return np.random.default_rng(5).uniform(0, 1, n)


def setup(nd=2, level=5, ti="ssp_rk3", kratio=100.0):
    (oc, oh), (c, h) = cht_setup(nd, level, kratio=kratio, source=source_field, ti=ti)
    return oc, oh, c, h


def oracle_J(oc, oh, snap, src, nsteps, dts):
    for s, q, t, dt, st in snap:
        s.set_lhs(q)
        s.set_time(t, dt, st)
    oh.set_source(src)
    J, _ = O.cht_objective(TR0, oc, oh, nsteps, dts)
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

    def fd(v, eps):
        return (oracle_J(oc, oh, osnap, S0 + eps * v, nsteps, dts)
                - oracle_J(oc, oh, osnap, S0 - eps * v, nsteps, dts)) / (2 * eps)

    for k in range(4):
        v = rng.uniform(-1, 1, n)
        v /= np.linalg.norm(v)
        gsnap.restore()
        _, dJ_t = A.cht_tangent(c, h, TR0, dts, v)
        dJ_a = float(grad @ v)
        # eps = 1e-6: the difference of two FP64 objectives carries a
        # rounding error of about 1e-16 |J| / eps, a few 1e-6 of dJ here;
        # eps = 1e-4 leaves 100x less rounding and an O(eps^2) truncation
        fd6, fd4 = fd(v, 1e-6), fd(v, 1e-4)
        print(f"direction {k}: adjoint {dJ_a:.12e} tangent {dJ_t:.12e} fd(1e-6) {fd6:.12e} fd(1e-4) {fd4:.12e}")
        assert abs(dJ_a - dJ_t) <= 1e-9 * abs(dJ_t), (k, dJ_a, dJ_t)
        assert abs(dJ_t - fd4) <= 1e-6 * abs(fd4), (k, dJ_t, fd4)
        assert abs(dJ_a - fd4) <= 1e-6 * abs(fd4), (k, dJ_a, fd4)
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
    for g in grads[1:]:
        np.testing.assert_allclose(g, grads[0], rtol=1e-11, atol=1e-13 * np.max(np.abs(grads[0])))
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
