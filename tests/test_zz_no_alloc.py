"""The allocation-free conservation sums (hom3g_solver_conservation_sums_noalloc)
against the allocating entry point, through the C ABI: two calls in a row
on the shared static scratch, the second on a smaller solver (fewer cells
and variables), each bit-identical to hom3g_solver_conservation_sums."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_sums_match_allocating_sums():
    import torch
    import hom3_amd as H
    from hom3_amd import _lib as L
    from hom3_amd import cases
    lib = L.lib()
    _, big = cases.tgv(H, 5)                      # 32^3 CNS, 5 variables
    g = H.Grid(2, [0.0, 0.0], [10.0, 10.0], 6)    # 64^2 Euler, 4 variables
    small = H.Solver(g, physics="euler", time_integration="ssp_rk3", CFL=0.5)
    H.append_bcs(small, H.Periodic())
    small.set_initial_condition(L.IC_ISENTROPIC_VORTEX, [0.0, 5.0, 5.0])
    small.initialize()
    big.solve(2)
    ctx = big.ctx
    out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for s in (big, small):
        ref = s.conservation_sums()
        rc = lib.hom3g_solver_conservation_sums_noalloc(ctx.h, s.id, C.c_void_p(out.data_ptr()))
        assert rc == 0, lib.hom3g_last_error()
        ctx.synchronize()
        got = out[: s.nvars].cpu().numpy()
        print(s.physics, "noalloc", got, "allocating", ref)
        np.testing.assert_array_equal(got, ref)
        assert np.all(np.isfinite(got)) and math.isfinite(float(ref[0]))
    assert lib.hom3g_solver_conservation_sums_noalloc(ctx.h, small.id, None) != 0
