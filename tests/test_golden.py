"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle reproduces them bit for bit (CPU), and the GPU
path matches them to the FP64 parity bar (GPU)."""
import os

import numpy as np
import pytest

from tests.golden import make_golden as MG
from tests.scenarios import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


@pytest.mark.parametrize("name", list(MG.CASES))
def test_oracle_reproduces_golden(name):
    ref = load(name)
    got = MG.make(name)
    np.testing.assert_array_equal(got["q0"], ref["q0"])
    np.testing.assert_array_equal(got["q1"], ref["q1"])
    assert got["dt1"] == ref["dt1"] and got["time"] == ref["time"]


GPU_CASES = {
    "vortex_muscl_rk3_16x16": lambda H: _vortex(H),
    "tgv_cns_rk3_8x8x8": lambda H: _tgv(H),
    "sod_euler_fe_8x8x8": lambda H: _sod(H),
    "cns_walls_fe_16x16": lambda H: _walls(H),
}


def _vortex(H):
    g = H.Grid(2, [-10, -10], [0, 0], 4)
    s = H.Solver(g, "euler", time_integration="ssp_rk3", recon="muscl", limiter="minmod", CFL=0.5)
    H.append_bcs(s, H.Periodic())
    return s


def _tgv(H):
    import math
    g = H.Grid(3, [0] * 3, [2 * math.pi] * 3, 3)
    s = H.Solver(g, "cns", time_integration="ssp_rk3", CFL=0.5, CFL_viscous=0.5, Minfty=0.1, ReInfty=1600.0,
                 Pr0=0.71)
    H.append_bcs(s, H.Periodic())
    return s


def _sod(H):
    g = H.Grid(3, [0] * 3, [1] * 3, 3)
    s = H.Solver(g, "euler", time_integration="euler_forward", CFL=0.2)
    H.append_bcs(s, H.euler.bc.Neumann())
    return s


def _walls(H):
    g = H.Grid(2, [0, 0], [1, 1], 4)
    s = H.Solver(g, "cns", time_integration="euler_forward", CFL=0.4, CFL_viscous=0.4, Minfty=0.1, ReInfty=1000.0)
    H.append_bcs(s, H.cns.bc.wall.AdiabaticNoSlip())
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(GPU_CASES))
def test_gpu_matches_golden(name):
    import hom3_amd as H
    ref = load(name)
    s = GPU_CASES[name](H)
    s.build()
    n = s.first_gc
    q = np.full((s.nvars, s.ncells), np.nan)
    q[:, :n] = ref["q0"]
    s.set_lhs(q)
    s.reset_time(0.0, float(ref["dt0"]), 0)
    s.solve(int(ref["nsteps"]))
    err = rel_l2(s.lhs[:, :n], ref["q1"], s.nd)
    assert np.max(err) <= 1e-12, err
    assert abs(s.time() - float(ref["time"])) <= 1e-14 * max(1.0, float(ref["time"]))
