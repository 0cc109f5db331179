"""Captured solver steps: solve(n) with n >= 3 runs one step eagerly, then
replays one captured step n - 1 times.  Every replay must reproduce the
eager single-step path bit for bit (state and time), with eager steps,
downloads and synchronisations between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNKS = (4, 1, 5, 3, 1, 6)


def _pair(make):
    g1, s1 = make()
    g2, s2 = make()
    return (g1, s1), (g2, s2)


def _check_replays(eager, graphed):
    done = 0
    for k in CHUNKS:
        graphed.solve(k)
        for _ in range(k):
            eager.solve(1)
        done += k
        a, b = eager.lhs, graphed.lhs
        assert np.all(np.isfinite(b)), done
        assert np.array_equal(a, b), (done, float(np.max(np.abs(a - b))))
        assert eager.time() == graphed.time(), done
    return done


def _vortex(ti="ssp_rk3", precision="fp64"):
    import hom3_amd as H
    from hom3_amd import _lib as L
    g = H.Grid(2, [0.0, 0.0], [10.0, 10.0], 6)
    s = H.Solver(g, physics="euler", time_integration=ti, recon="muscl", limiter="minmod",
                 precision=precision, gamma=1.4, CFL=0.5)
    H.append_bcs(s, H.Periodic())
    s.set_initial_condition(L.IC_ISENTROPIC_VORTEX, [0.0, 5.0, 5.0])
    s.initialize()
    return g, s


@pytest.mark.parametrize("ti", ["ssp_rk3", "runge_kutta_2"])
def test_brick_vortex_replays_match_eager(ti):
    (_, a), (_, b) = _pair(lambda: _vortex(ti))
    assert b.kernel() == "brick"
    _check_replays(a, b)


def test_brick_tgv_fp32_replays_match_eager():
    import hom3_amd as H
    from hom3_amd import cases
    (_, a), (_, b) = _pair(lambda: cases.tgv(H, 4, precision="fp32"))
    assert b.kernel() == "brick"
    _check_replays(a, b)


def test_generic_blast_replays_match_eager():
    import hom3_amd as H
    from hom3_amd import cases

    def make():
        g, s, _ = cases.blast(H, base=4, extra=1)
        return g, s
    (_, a), (_, b) = _pair(make)
    assert b.kernel() == "generic"
    _check_replays(a, b)


def test_replay_follows_new_state_and_time():
    """Uploading a state or resetting the time changes device data only: the
    same captured step keeps stepping from the new values."""
    (_, a), (_, b) = _pair(lambda: _vortex())
    b.solve(4)
    q0 = a.lhs
    b.set_lhs(q0)
    b.reset_time(0.0, 0.0, 0)
    a.reset_time(0.0, 0.0, 0)
    _check_replays(a, b)


def test_forced_dt_then_replays():
    """A forced time step is device data read by the eager first step of
    solve(n); the replays after it take the CFL step again."""
    (_, a), (_, b) = _pair(lambda: _vortex())
    a.dt(1e-3)
    b.dt(1e-3)
    _check_replays(a, b)
    assert a.time() > 1e-3
