"""The brick reverse kernel's face-flux adjoint (brick_face_rev,
hom3_amd/csrc/face_rev.hpp) against central differences of the forward
face flux the brick kernel evaluates (brick_prims + ausm_flux_fast +
viscous_flux_fast), in long double, for every input component on random
3D and 2D faces along each direction (Mach 0.05 to 3: both upwind sides,
sub- and supersonic branches).  Host-only: tests/cpp/face_rev_check.cpp
is compiled with g++ against the library's device headers and a host
stand-in for the HIP runtime header (tests/cpp/host_stub)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_face_reverse_matches_central_differences(tmp_path):
    exe = str(tmp_path / "face_rev_check")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(HERE, "cpp", "host_stub"),
           "-I", os.path.join(ROOT, "hom3_amd", "csrc"), os.path.join(HERE, "cpp", "face_rev_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:]
    worst = float(r.stdout.strip().split()[-1])
    assert worst < 1e-7
