"""The C ABI: libhom3g.so loads, exports every entry point include/hom3g.h
declares, the ctypes structs match the C layout, and calls fail loudly (with
an error status, not a crash or a silent fallback) when no GPU is present.
CPU only: no compute call is made here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hom3g.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(hom3g_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    for must in ("hom3g_create", "hom3g_solver_solve", "hom3g_solver_build", "hom3g_couple_cht_step",
                 "hom3g_grid_neighbors", "hom3g_solver_topology", "hom3g_last_error"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from hom3_amd import _lib
    lib = _lib.lib()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(hom3g_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_struct_layout_matches_c():
    from hom3_amd import _lib
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "hom3g.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(hom3g_physics), offsetof(hom3g_physics, gamma),
         offsetof(hom3g_physics, time_end), offsetof(hom3g_physics, max_cells),
         offsetof(hom3g_physics, precision), offsetof(hom3g_physics, kernel));
  printf("%zu %zu %zu\n", sizeof(hom3g_geometry), offsetof(hom3g_geometry, a), offsetof(hom3g_geometry, b));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    P = _lib.Physics
    G = _lib.Geometry
    assert lines[0].split() == [str(v) for v in (C.sizeof(P), P.gamma.offset, P.time_end.offset,
                                                 P.max_cells.offset, P.precision.offset, P.kernel.offset)]
    assert lines[1].split() == [str(v) for v in (C.sizeof(G), G.a.offset, G.b.offset)]


def test_oracle_and_abi_physics_prefix_identical():
    import oracle as O
    from hom3_amd import _lib
    for name, _t in O.PhysDesc._fields_:
        assert getattr(O.PhysDesc, name).offset == getattr(_lib.Physics, name).offset


def test_create_without_gpu_fails_loudly():
    """No silent CPU path: without a HIP device the context cannot be made."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from hom3_amd import _lib
    h = C.c_void_p()
    rc = _lib.lib().hom3g_create(C.byref(h), 0)
    assert rc != 0
    assert _lib.lib().hom3g_last_error()
    import hom3_amd
    with pytest.raises(_lib.Hom3gError):
        hom3_amd.Context(0)


def test_null_context_is_an_error_not_a_crash():
    from hom3_amd import _lib
    L = _lib.lib()
    assert L.hom3g_synchronize(None) == -1
    n = C.c_int64()
    assert L.hom3g_grid_size(None, 0, C.byref(n)) == -1
    assert b"null context" in L.hom3g_last_error()
