"""The C++ facade (include/hom3g/fv_solver.hpp) driving the library from a
hom3-style program: tests/cpp/facade_demo.cpp, built by __graft_entry__.build(),
runs the 3D Sod tube of euler_fv_solver_test.cpp:126-156 (level 3, forward
Euler, Neumann boundaries) and prints its state; compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests.scenarios import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "facade_demo")


def test_facade_demo_matches_oracle():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/facade_demo is not built (run __graft_entry__.build())")
    steps = 4
    r = subprocess.run([EXE, str(steps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    n, nint, nv, t, dt = lines[0].split()
    n, nint, nv = int(n), int(nint), int(nv)
    q = np.array([float(x) for x in lines[1:1 + nv * nint]]).reshape(nv, nint)
    g = O.Grid(3, [0.0] * 3, [1.0] * 3).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_EULER, nd=3, cfl=0.2, time_end=0.2))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    s.set_ic(O.IC_SHOCK_TUBE, [0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1])
    s.initialize()
    s.solve(steps)
    assert s.ncells == n and s.first_gc == nint
    err = rel_l2(q, s.get("lhs")[:, :nint], 3)
    assert np.max(err) <= 1e-12, err
    assert abs(float(t) - s.time()) <= 1e-13 * s.time()
