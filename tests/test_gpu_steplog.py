"""JSON-lines step log: one record per output step with step, time, dt, the
device conservation sums and the bad-cell count; on the periodic TGV the
mass and energy sums stay constant to round-off."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_step_log_records_and_conservation(tmp_path):
    import hom3_amd as H
    from hom3_amd import cases
    from hom3_amd.steplog import StepLog
    g, s = cases.tgv(H, 4)
    path = tmp_path / "tgv.jsonl"
    with StepLog(str(path)) as log:
        log(s)
        for _ in range(4):
            s.solve(2)
            log(s)
    recs = [json.loads(line) for line in path.read_text().splitlines()]
    assert [r["step"] for r in recs] == [0, 2, 4, 6, 8]
    assert all(r["bad_cells"] == 0 for r in recs)
    assert all(b["time"] > a["time"] and b["dt"] > 0 for a, b in zip(recs, recs[1:]))
    sums = np.array([r["sums"] for r in recs])
    for v in (3, 4):  # rho, rho E: periodic box, conserved
        assert np.max(np.abs(sums[:, v] - sums[0, v])) <= 1e-12 * abs(sums[0, v]), sums[:, v]
