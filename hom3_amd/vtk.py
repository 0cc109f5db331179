"""Legacy VTK output of a solver's device state (io/vtk.hpp:168-405,
solver.hpp:329-376 write_domain).

Same file layout as the reference's binary writer: an UNSTRUCTURED_GRID of
pixels (2D, type 8) or voxels (3D, type 11) for the internal cells followed
by the ghost cells, unique vertices sorted lexicographically (x, then y,
then z), big-endian int32 ids and float32 values, and the cell streams
locallCellIds, gridNodeIds, local_nghbrs_<pos>, ghost_cell, bcIdx, the
conserved variables (rho_u<d>, rho, rho_E or T) and the physics output
(a, p, u_<d>, M, e, T for Euler/Navier-Stokes).  Invalid ids are written as
-1.  The state is downloaded through the C ABI; nothing is computed on the
host but the output quantities.
"""
import numpy as np

from ._lib import INVALID_ID


def _ids(a):
    a = np.asarray(a, np.int64)
    return np.where((a == INVALID_ID) | (a < -1), -1, a).astype(">i4")


def cell_vertices(xc, length, nd):
    """(ncells * 2^nd, nd) vertices, pixel/voxel order (bit d of the vertex
    index selects the upper side along d)."""
    n = xc.shape[1]
    nv = 1 << nd
    out = np.empty((n, nv, nd))
    for k in range(nv):
        for d in range(nd):
            out[:, k, d] = xc[d] + (0.5 if (k >> d) & 1 else -0.5) * length
    return out.reshape(n * nv, nd)


def write_domain(path, solver, binary=True):
    """Write the solver's cells (internal, then ghosts) and state to
    ``path`` (".vtk" appended when missing)."""
    if not path.endswith(".vtk"):
        path += ".vtk"
    nd = solver.nd
    t = solver.topology()
    q = solver.lhs
    n = q.shape[1]
    n_int = solver.first_gc
    xc = t["x_center"]
    # ghost cells carry the length of their boundary cell (solver.hpp:761-773)
    length = t["length"][0]
    nb = t["neighbors"]
    verts = cell_vertices(xc, length, nd)
    scale = max(float(np.max(np.abs(verts))), 1.0)
    key = np.round(verts / (scale * 1e-12)).astype(np.int64)
    order = np.lexsort(tuple(key[:, d] for d in reversed(range(nd))))
    uniq_mask = np.ones(len(order), bool)
    ks = key[order]
    uniq_mask[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    vid_sorted = np.cumsum(uniq_mask) - 1
    vertex_id = np.empty(len(order), np.int64)
    vertex_id[order] = vid_sorted
    points = verts[order][uniq_mask]
    nvert = 1 << nd
    cells = np.empty((n, nvert + 1), np.int64)
    cells[:, 0] = nvert
    cells[:, 1:] = vertex_id.reshape(n, nvert)

    streams = []  # (name, kind, values)
    streams.append(("locallCellIds", "int", np.arange(n)))
    streams.append(("gridNodeIds", "int", t["node_idx"][0]))
    for p in range(2 * nd):
        streams.append((f"local_nghbrs_{p}", "int", nb[p]))
    streams.append(("ghost_cell", "int", np.where(np.arange(n) >= n_int, 1, -1)))
    streams.append(("bcIdx", "int", t["bc_idx"][0].astype(np.int64)))
    if solver.nvars == 1:
        streams.append(("T", "float", q[0]))
    else:
        names = [f"rho_u{d}" for d in range(nd)] + ["rho", "rho_E"]
        for v, nm in enumerate(names):
            streams.append((nm, "float", q[v]))
        gamma = solver.desc.gamma
        rho = q[nd]
        u = q[:nd] / rho
        p = (gamma - 1.0) * (q[nd + 1] - 0.5 * rho * np.sum(u * u, axis=0))
        a = np.sqrt(gamma * p / rho)
        streams.append(("a", "float", a))
        streams.append(("p", "float", p))
        for d in range(nd):
            streams.append((f"u_{d}" if nd > 1 else "u", "float", u[d]))
        streams.append(("M", "float", np.sqrt(np.sum(u * u, axis=0)) / a))
        streams.append(("e", "float", (q[nd + 1] - 0.5 * rho * np.sum(u * u, axis=0)) / rho))
        streams.append(("T", "float", gamma * p / rho))

    with open(path, "wb") as f:
        def text(s):
            f.write(s.encode())

        text("# vtk DataFile Version 2.0\nHOM3 legacy vtk Output \n")
        text("BINARY\n" if binary else "ASCII\n")
        text("DATASET UNSTRUCTURED_GRID\n")
        pts = np.zeros((len(points), 3))
        pts[:, :nd] = points
        text(f"POINTS {len(points)} FLOAT\n")
        _write(f, pts.astype(">f4"), binary)
        text(f"CELLS {n} {n * (nvert + 1)}\n")
        _write(f, cells.astype(">i4"), binary)
        text(f"CELL_TYPES {n}\n")
        _write(f, np.full(n, 8 if nd == 2 else 11, ">i4"), binary)
        text(f"CELL_DATA {n}\n")
        for name, kind, vals in streams:
            text(f"SCALARS {name} {kind}\nLOOKUP_TABLE default\n")
            arr = _ids(vals) if kind == "int" else np.asarray(vals, ">f4")
            _write(f, arr, binary)
    return path


def _write(f, arr, binary):
    if binary:
        f.write(np.ascontiguousarray(arr).tobytes())
        f.write(b"\n")
    else:
        flat = np.asarray(arr).astype(np.float64 if arr.dtype.kind == "f" else np.int64).ravel()
        f.write((" ".join(repr(float(x)) if arr.dtype.kind == "f" else str(int(x)) for x in flat) + "\n").encode())


def read_cell_data(path):
    """Minimal reader of the binary files written above (for tests):
    returns (points, cells, {name: values})."""
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    for _ in range(4):
        line()
    hdr = line().split()
    npts = int(hdr[1])
    pts = np.frombuffer(data, ">f4", npts * 3, pos).reshape(npts, 3)
    pos += npts * 12 + 1
    hdr = line().split()
    n, tot = int(hdr[1]), int(hdr[2])
    cells = np.frombuffer(data, ">i4", tot, pos).reshape(n, -1)
    pos += tot * 4 + 1
    line()
    pos += n * 4 + 1
    line()
    out = {}
    while pos < len(data):
        hdr = line().split()
        name, kind = hdr[1], hdr[2]
        line()
        arr = np.frombuffer(data, ">i4" if kind == "int" else ">f4", n, pos)
        pos += n * 4 + 1
        out[name] = arr.astype(np.int64 if kind == "int" else np.float64)
    return pts, cells, out
