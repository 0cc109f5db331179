"""The oracle's long-double build (the checker for central differences at
small steps) against its FP64 build: same objective to rounding on the
coupled CHT case, and central differences at eps = 1e-6 and 1e-4 that agree
with each other far below what FP64 reaches at eps = 1e-6."""
import numpy as np

import oracle as O


def _cht(Om, level=4, nd=2):
    # the oracle side of the coupled CHT case (flow past a heated block)
    from tests.test_gpu_cht import cht_setup_oracle
    return cht_setup_oracle(Om, nd=nd, level=level, kratio=100.0,
                            source=lambda n: np.random.default_rng(5).uniform(0, 1, n), ti="ssp_rk3")


def _snap(*ss):
    return [(s, s.get("lhs").copy(), s.time(), s.dt(), s.step()) for s in ss]


def _J(Om, oc, oh, snap, src, nsteps, dts):
    for s, q, t, dt, st in snap:
        s.set_lhs(q)
        s.set_time(t, dt, st)
    oh.set_source(src)
    return Om.cht_objective(0.01, oc, oh, nsteps, dts)[0]


def test_long_double_oracle_matches_and_differentiates_cleanly():
    OL = O.long_double()
    oc, oh = _cht(O)
    ocl, ohl = _cht(OL)
    # identical topology, and the long-double solver started from the FP64 state
    np.testing.assert_array_equal(oc.get("neighbors"), ocl.get("neighbors"))
    snap = _snap(oc, oh)
    lsnap = [(sl,) + tuple(rest) for sl, (_, *rest) in zip((ocl, ohl), snap)]
    n = oh.first_gc
    S0 = np.random.default_rng(5).uniform(0, 1, n)
    nsteps = 8
    J64, dts = O.cht_objective(0.01, oc, oh, nsteps)
    J64 = _J(O, oc, oh, snap, S0, nsteps, dts)
    JL = _J(OL, ocl, ohl, lsnap, S0, nsteps, dts)
    assert abs(JL - J64) <= 1e-9 * abs(J64), (JL, J64)
    v = np.random.default_rng(11).uniform(-1, 1, n)
    v /= np.linalg.norm(v)

    def fd(Om, a, b, sn, eps):
        return (_J(Om, a, b, sn, S0 + eps * v, nsteps, dts) - _J(Om, a, b, sn, S0 - eps * v, nsteps, dts)) / (2 * eps)

    l6, l4 = fd(OL, ocl, ohl, lsnap, 1e-6), fd(OL, ocl, ohl, lsnap, 1e-4)
    d6, d4 = fd(O, oc, oh, snap, 1e-6), fd(O, oc, oh, snap, 1e-4)
    print(f"long double: fd(1e-6) {l6:.15e} fd(1e-4) {l4:.15e}; fp64: fd(1e-6) {d6:.15e} fd(1e-4) {d4:.15e}")
    # the eps^2 truncation at 1e-4 and the long-double rounding at 1e-6 are
    # both far below 1e-7 of dJ; FP64 at 1e-4 agrees to its own rounding
    assert abs(l6 - l4) <= 1e-7 * abs(l6), (l6, l4)
    assert abs(d4 - l6) <= 1e-6 * abs(l6), (d4, l6)
