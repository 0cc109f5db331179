"""Edge cases at the boundary: a cell count above maxNoCells or the brick
kernel asked for where it does not apply are errors with their codes
(HOM3G_ERR_CAPACITY, HOM3G_ERR_ARG), not crashes; a zero-step solve
changes nothing and a negative one is refused; and the tiled brick kernel
is taken only when the solver's cells form whole Morton bricks (the channel
at level 4 has a fluid 8 cells wide starting at cell 4, which is not
brick-aligned; at level 5 it is)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_CAPACITY = -1, -5


def _blast(H, level, refine=False, **kw):
    from hom3_amd import _lib as L, cases
    g = H.Grid(3, [0.0] * 3, [1.0] * 3, level)
    if refine:
        g.refine_region(level + 1, H.Sphere([0.5, 0.5, 0.5], 0.2), 0.05)
        g.balance()
    s = H.Solver(g, physics="euler", time_integration="ssp_rk3", CFL=0.2, **kw)
    H.append_bcs(s, H.euler.bc.Neumann())
    s.set_initial_condition(L.IC_SPHERE, cases.BLAST_IC)
    return g, s


def test_max_no_cells_exceeded_is_an_error():
    import hom3_amd as H
    from hom3_amd._lib import Hom3gError
    _, s = _blast(H, 4, maxNoCells=1000)  # 4,096 cells and their ghosts
    with pytest.raises(Hom3gError) as e:
        s.initialize()
    assert e.value.code == ERR_CAPACITY
    # the same grid builds with room for it
    _, s2 = _blast(H, 4, maxNoCells=8192)
    s2.initialize()
    assert s2.first_gc == 4096


def test_brick_kernel_on_a_refined_grid_is_an_error():
    import hom3_amd as H
    from hom3_amd._lib import Hom3gError
    _, s = _blast(H, 3, refine=True, kernel="brick")
    with pytest.raises(Hom3gError) as e:
        s.initialize()
    assert e.value.code == ERR_ARG and "brick" in str(e.value)
    _, s2 = _blast(H, 3, refine=True)  # auto: the table-driven kernel
    s2.initialize()
    assert s2.kernel() == "generic"


def test_solve_zero_steps_changes_nothing():
    import hom3_amd as H
    from hom3_amd import cases
    g, s = cases.tgv(H, 4)
    s.solve(2)
    q0, t0, n0 = s.lhs.copy(), s.time(), s.step()
    s.solve(0)
    np.testing.assert_array_equal(s.lhs, q0)
    assert s.time() == t0 and s.step() == n0
    from hom3_amd._lib import Hom3gError
    with pytest.raises(Hom3gError) as e:
        s.solve(-1)
    assert e.value.code == ERR_ARG
    np.testing.assert_array_equal(s.lhs, q0)


@pytest.mark.parametrize("level,kernel", [(4, "generic"), (5, "brick")])
def test_channel_fluid_takes_tiled_bricks_only_when_brick_aligned(level, kernel):
    import hom3_amd as H
    from hom3_amd import cases
    g, c, h, _ = cases.channel_cht(H, level=level)
    assert c.kernel() == kernel
    assert h.kernel() == "generic"
