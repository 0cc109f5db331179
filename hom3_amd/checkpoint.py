"""HDF5 checkpoint / restart of solver states (libhom3g_io.so, include/hom3g_io.h).

``write(path, *solvers)`` stores every solver's cells().lhs and clock in the
group "/solver<k>"; ``read(path, *solvers)`` restores them after checking
that the topology matches.  Used to resume a run (the reference's unused
"restart" property, solver.hpp:88-95)."""
import ctypes as C
import os

import numpy as np

from . import _lib

IO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhom3g_io.so")
_io = None


def lib():
    global _io
    if _io is None:
        _lib.lib()  # libhom3g first: the io library resolves against it
        if not os.path.exists(IO_PATH):
            raise RuntimeError(f"{IO_PATH} is not built (HDF5 missing at build time?)")
        L = C.CDLL(IO_PATH)
        for name in ("hom3g_io_write_checkpoint", "hom3g_io_read_checkpoint"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_char_p]
        L.hom3g_io_last_error.restype = C.c_char_p
        L.hom3g_io_last_error.argtypes = []
        _io = L
    return _io


def _call(fn, path, solvers):
    ids = np.ascontiguousarray([s.id for s in solvers], np.int32)
    rc = fn(solvers[0].ctx.h, ids.ctypes.data_as(C.POINTER(C.c_int32)), ids.size, os.fsencode(path))
    if rc != 0:
        raise _lib.Hom3gError(rc, lib().hom3g_io_last_error().decode())


def write(path, *solvers):
    _call(lib().hom3g_io_write_checkpoint, path, solvers)


def read(path, *solvers):
    _call(lib().hom3g_io_read_checkpoint, path, solvers)
