"""JSON-lines step log for solver runs (the reference prints step, time and
dt to std::cerr, utilities.hpp:83-95; this writes them as one JSON object
per line, with the device conservation sums and the bad-cell count).

Use as the ``on_output`` callback of ``fv.run_solver`` or call it after any
``solve``::

    log = StepLog("run.jsonl")
    run_solver(grid, solver, 100, output_interval=10, on_output=log)
    log.close()
"""
import json
import time


class StepLog:
    def __init__(self, path, sums=True, check=True):
        self.f = open(path, "w")
        self.sums = sums
        self.check = check
        self.t0 = time.perf_counter()

    def __call__(self, solver):
        rec = dict(step=int(solver.step()), time=float(solver.time()), dt=float(solver.dt()),
                   wall_s=round(time.perf_counter() - self.t0, 6))
        if self.sums:
            rec["sums"] = [float(x) for x in solver.conservation_sums()]
        if self.check:
            rec["bad_cells"] = int(solver.check_state())
        self.f.write(json.dumps(rec) + "\n")
        self.f.flush()
        return rec

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
