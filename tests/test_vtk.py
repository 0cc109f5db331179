"""VTK output (io/vtk.hpp:168-405, solver.hpp:329-376): file structure,
unique vertices and cell streams, on topology and state taken from the
oracle (the writer only formats what the solver interface returns)."""
import numpy as np

import oracle as O
from hom3_amd import vtk


class _FromOracle:
    """The part of the solver interface the writer reads."""

    def __init__(self, o, gamma=1.4):
        self.o = o
        self.nd = o.nd
        self.nvars = o.nvars
        self.first_gc = o.first_gc

        class D:
            pass
        self.desc = D()
        self.desc.gamma = gamma

    def topology(self):
        return {k: self.o.get(k) for k in ("neighbors", "node_idx", "bc_idx", "x_center", "length", "distances")}

    @property
    def lhs(self):
        return self.o.get("lhs")


def _solver(nd, level):
    g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(level)
    s = O.Solver(g, O.phys(physics=O.PHYS_EULER, nd=nd))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * (2 * nd))
    s.set_ic(O.IC_FOUR_QUADRANT if nd == 2 else O.IC_CONSTANT_PV,
             [0.5, 0.5, 0.1, 0.2, 1.0, 1.0, -0.3, 0.1, 0.5, 0.4, 0.0, 0.0, 2.0, 1.5, 0.2, -0.2, 0.8, 0.9]
             if nd == 2 else [0.1, 0.2, 0.3, 1.0, 1.0])
    s.initialize()
    return s


def test_vtk_roundtrip_2d(tmp_path):
    s = _solver(2, 3)
    f = vtk.write_domain(str(tmp_path / "dom"), _FromOracle(s))
    pts, cells, data = vtk.read_cell_data(f)
    n, n_int = s.ncells, s.first_gc
    assert cells.shape == (n, 5) and np.all(cells[:, 0] == 4)
    # 8x8 internal cells (9x9 vertices) + one strip of 8 ghosts per side,
    # each adding an outer line of 9 vertices
    assert len(pts) == 9 * 9 + 4 * 9
    # points sorted by x, then y
    assert np.all(np.diff(pts[:, 0]) >= 0)
    np.testing.assert_array_equal(data["locallCellIds"], np.arange(n))
    np.testing.assert_array_equal(data["ghost_cell"], np.where(np.arange(n) >= n_int, 1, -1))
    nb = s.get("neighbors")
    for p in range(4):
        np.testing.assert_array_equal(data[f"local_nghbrs_{p}"], np.where(nb[p] == O.INVALID, -1, nb[p]))
    q = s.get("lhs")
    np.testing.assert_array_equal(data["rho"], q[2].astype(np.float32))
    # each cell's vertices span its own box
    xc = s.get("x_center")
    h = s.get("length")[0]
    for c in (0, 17, n_int, n - 1):
        v = pts[cells[c, 1:]]
        np.testing.assert_allclose(v[:, :2].min(0), xc[:, c] - h[c] / 2, atol=1e-6)
        np.testing.assert_allclose(v[:, :2].max(0), xc[:, c] + h[c] / 2, atol=1e-6)


def test_vtk_3d_voxels(tmp_path):
    s = _solver(3, 2)
    f = vtk.write_domain(str(tmp_path / "dom3"), _FromOracle(s))
    pts, cells, data = vtk.read_cell_data(f)
    assert cells.shape == (s.ncells, 9)
    assert set(data) >= {"rho_u0", "rho_u1", "rho_u2", "rho", "rho_E", "a", "p", "u_0", "M", "e", "T", "bcIdx"}
    np.testing.assert_allclose(data["p"][: s.first_gc], 1.0, rtol=1e-6)
