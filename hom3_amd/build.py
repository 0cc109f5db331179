"""Build libhom3g.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m hom3_amd.build  (or hom3_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhom3g.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["abi.hip", "topology.hip", "stepper.hip", "brick.hip", "batch.hip", "adjoint.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HOM3G_ARCH", "gfx950")

FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-Wl,-soname,libhom3g.so", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_io(verbose)
    return OUT


HDF5_ROOT = os.environ.get("HOM3G_HDF5_ROOT", "/opt/conda")
IO_OUT = os.path.join(HERE, "libhom3g_io.so")


def build_io(verbose=True):
    """The HDF5 checkpoint library (host C++ over the C ABI); skipped when
    no HDF5 installation is found."""
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        if verbose:
            print(f"HDF5 not found under {HDF5_ROOT}: libhom3g_io.so not built")
        return None
    tmp = IO_OUT + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall",
           "-I", os.path.join(HERE, "..", "include"), "-I", os.path.join(HDF5_ROOT, "include"),
           os.path.join(HERE, "io", "checkpoint.cpp"), "-L", HERE, "-lhom3g",
           "-L", os.path.join(HDF5_ROOT, "lib"), "-lhdf5_hl", "-lhdf5",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libhom3g_io build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, IO_OUT)
    if verbose:
        print(f"built {IO_OUT}")
    return IO_OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
