"""The HDF5 checkpoint library loads next to libhom3g and exports what
include/hom3g_io.h declares (no GPU needed)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_io_library_exports_header_symbols():
    from hom3_amd import checkpoint
    if not os.path.exists(checkpoint.IO_PATH):
        pytest.skip("HDF5 not available at build time")
    hdr = open(os.path.join(ROOT, "include", "hom3g_io.h")).read()
    names = set(re.findall(r"\b(hom3g_io_\w+)\s*\(", hdr))
    assert names == {"hom3g_io_write_checkpoint", "hom3g_io_read_checkpoint", "hom3g_io_last_error"}
    out = subprocess.run(["nm", "-D", "--defined-only", checkpoint.IO_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert names <= exported
    checkpoint.lib()  # loads against the already loaded libhom3g
