"""The brick kernel's closed-form halo enumeration (depth-1 halos) covers
exactly the positions of its compile-time halo list, face halos first, for
3D/2D Navier-Stokes and Euler bricks.  Host-only: compiles
tests/cpp/brick_halo_check.hip with hipcc and runs it on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not available")
def test_closed_form_halo_matches_list(tmp_path):
    exe = str(tmp_path / "brick_halo_check")
    cmd = [HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "hom3_amd", "csrc"),
           os.path.join(HERE, "cpp", "brick_halo_check.hip"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("OK") == 4
