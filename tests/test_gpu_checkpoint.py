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
