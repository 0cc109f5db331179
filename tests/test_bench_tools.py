"""Host logic of bench.py's measurement tools: the HBM traffic per stage
launch from rocprofv3 counter files (calibrated by the copy of the same
run), with the provenance of the kernel sources it was measured on."""
import csv
import json

import bench


def _write_pmc(path, counter, rows):
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for disp, name, val in rows:
            w.writerow(dict(Dispatch_Id=disp, Kernel_Name=name, Counter_Name=counter, Counter_Value=val))


def test_write_traffic_scales_by_the_copy(tmp_path):
    stage = "void hom3g::(anonymous namespace)::k_brick<3, 1, false, double, double, double, false>(args)"
    copy = "hom3g::(anonymous namespace)::k_copy(double4 const*, double4*, long long)"
    gib_kib = float(1 << 30) / 1024
    # FETCH counts half of a streaming read (factor 2), WRITE all of it
    fetch = [(i, stage, 1000.0 + i) for i in range(8)] + [(100, copy, gib_kib / 2)]
    write = [(i, stage, 500.0) for i in range(8)] + [(100, copy, gib_kib)]
    # a second counter row of the same dispatch (another XCD) adds up
    fetch.append((0, stage, 0.0))
    fa, wa = tmp_path / "f.csv", tmp_path / "w.csv"
    _write_pmc(fa, "FETCH_SIZE", fetch)
    _write_pmc(wa, "WRITE_SIZE", write)
    out = tmp_path / "t.json"
    d = bench.write_traffic(str(fa), str(wa), level=8, out_path=str(out))
    # the first three launches (first step) are skipped: mean of 1003..1007 KiB x 2 + 500 KiB
    expect = (sum(1000.0 + i for i in range(3, 8)) / 5) * 2 * 1024 + 500.0 * 1024
    assert abs(d["traffic_bytes_per_launch"] - expect) <= 1
    assert abs(d["fetch_factor"] - 2.0) < 1e-9 and abs(d["write_factor"] - 1.0) < 1e-9
    assert d["stage_kernel_sha16"] == bench.stage_kernel_sha16() and d["launches"] == 5
    assert json.load(open(out)) == d
