"""HDF5 checkpoint / restart: a run resumed from a checkpoint reproduces the
uninterrupted run bit for bit; a checkpoint of another topology is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(H, level):
    from hom3_amd import cases
    return cases.tgv(H, level)


def test_restart_reproduces_run(tmp_path):
    import hom3_amd as H
    from hom3_amd import checkpoint
    g, s = _solver(H, 4)
    s.solve(3)
    path = str(tmp_path / "tgv.h5")
    checkpoint.write(path, s)
    s.solve(4)
    qa, ta = s.lhs.copy(), (s.time(), s.dt(), s.step())
    g2, s2 = _solver(H, 4)
    checkpoint.read(path, s2)
    assert s2.step() == 3
    s2.solve(4)
    np.testing.assert_array_equal(s2.lhs[:, :s2.first_gc], qa[:, :s.first_gc])
    assert (s2.time(), s2.dt(), s2.step()) == ta
    g3, s3 = _solver(H, 3)
    with pytest.raises(Exception):
        checkpoint.read(path, s3)


def _h5():
    """ctypes access to the HDF5 high-level library (reading the file as the
    reference's HDF5File would: H5LTread_dataset, H5LTget_attribute_*)."""
    import ctypes as C
    L = C.CDLL("/opt/conda/lib/libhdf5_hl.so")
    H = C.CDLL("/opt/conda/lib/libhdf5.so")
    H.H5open()
    H.H5Fopen.restype = C.c_int64
    H.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    H.H5Fclose.argtypes = [C.c_int64]
    L.H5LTget_attribute_long_long.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.POINTER(C.c_longlong)]
    L.H5LTget_attribute_int.argtypes = [C.c_int64, C.c_char_p, C.c_char_p, C.POINTER(C.c_int)]
    L.H5LTread_dataset_double.argtypes = [C.c_int64, C.c_char_p, C.POINTER(C.c_double)]
    return C, H, L


def test_restart_in_the_reference_layout_reproduces_the_oracle_run(tmp_path):
    """An oracle run of 7 steps against a GPU run of 3 steps, an HDF5
    checkpoint in the reference's HDF5File layout (io/hdf5_file.hpp:144-158:
    rank-1 datasets with no_rows, no_cols, order), a restart into a fresh
    solver and 4 more steps; the file read back with the HDF5 high-level
    calls the reference uses holds the checkpointed state."""
    from hom3_amd import checkpoint
    from tests.scenarios import build_gpu, build_oracle, rel_l2
    from tests.test_gpu_parity import STEP_SCENARIOS
    sc = STEP_SCENARIOS["cns2d_isothermal_walls_rk2"]()
    _, o = build_oracle(sc)
    _, s = build_gpu(sc)
    s.set_lhs(o.get("lhs"))
    s.reset_time(o.time(), o.dt(), o.step())
    s.solve(3)
    q3 = s.lhs.copy()
    path = str(tmp_path / "cns.hdf5")
    checkpoint.write(path, s)
    _, s2 = build_gpu(sc)
    checkpoint.read(path, s2)
    assert s2.step() == 3
    s2.solve(4)
    o.solve(7)
    n = o.first_gc
    err = rel_l2(s2.lhs[:, :n], o.get("lhs")[:, :n], 2)
    assert np.max(err) <= 1e-12, err
    assert s2.step() == 7 and abs(s2.time() - o.time()) <= 1e-13 * o.time()
    C, H, L = _h5()
    f = H.H5Fopen(path.encode(), 0, 0)
    assert f >= 0
    try:
        v = C.c_longlong()
        i = C.c_int()
        assert L.H5LTget_attribute_long_long(f, b"/", b"no_solvers", C.byref(v)) >= 0 and v.value == 1
        assert L.H5LTget_attribute_long_long(f, b"/solver0/lhs", b"no_rows", C.byref(v)) >= 0
        assert v.value == s.ncells
        assert L.H5LTget_attribute_long_long(f, b"/solver0/lhs", b"no_cols", C.byref(v)) >= 0 and v.value == 4
        assert L.H5LTget_attribute_int(f, b"/solver0/lhs", b"order", C.byref(i)) >= 0 and i.value == 1
        assert L.H5LTget_attribute_long_long(f, b"/solver0", b"step", C.byref(v)) >= 0 and v.value == 3
        buf = np.empty(4 * s.ncells)
        assert L.H5LTread_dataset_double(f, b"/solver0/lhs", buf.ctypes.data_as(C.POINTER(C.c_double))) >= 0
        np.testing.assert_array_equal(buf.reshape(4, s.ncells), q3)
    finally:
        H.H5Fclose(f)
