"""Build libgatt_dns.so for gfx950 with hipcc (in-tree, no JIT cache)."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libgatt_dns.so")
SOURCES = ["line_solve_x_xg.hip", "line_solve_x_wsp.hip", "line_solve_y_wsp.hip", "line_solve_z_wsp.hip",
           "line_solve_x_ws.hip", "line_solve_y_ws.hip", "line_solve_z_ws.hip", "line_solve_x_lg.hip",
           "line_solve_y_lg.hip", "line_solve_z_lg.hip", "line_solve_multi.hip", "pointwise.hip", "boundary.hip", "lst.hip", "line_solve_t.hip", "adjoint.hip", "rhs_line.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GATT_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         # packed FP32 instructions are disabled for every HIP object
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # every expression is rounded as written (no FMA contraction), like the
         # reference's statement-by-statement evaluation
         "-ffp-contract=off",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


FLOWIO_LIB = os.path.join(OUT_DIR, "libgatt_flowio.so")
HDF5_DIR = os.environ.get("GATT_HDF5_DIR", "/opt/conda")


def build_flowio(verbose=True):
    """Host library of the MAT 7.3 flow files (flow_io.cpp, libhdf5).  Skipped with a
    message when the HDF5 headers are not found; returns the library path or None."""
    inc = os.path.join(HDF5_DIR, "include")
    libdir = os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print("flow-file library skipped: no hdf5.h under", inc)
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    cxx = os.environ.get("CXX", shutil.which("g++") or "g++")
    tmp = FLOWIO_LIB + ".tmp"
    # the C++ runtime is linked in statically: the HDF5 directory on the run path must not
    # supply the process's libstdc++ (an older one there would break the HIP runtime)
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-static-libstdc++", "-static-libgcc", "-I", inc,
           os.path.join(SRC, "flow_io.cpp"), "-L", libdir, "-lhdf5", "-lz", "-Wl,-rpath," + libdir, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flow-file library build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, FLOWIO_LIB)
    if verbose:
        print("built", FLOWIO_LIB)
    return FLOWIO_LIB


def _stale(src, obj):
    """An object is out of date when any source or header it may include is newer."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(SRC, src)] + [os.path.join(SRC, h) for h in os.listdir(SRC) if h.endswith(".h")]
    inc = os.path.join(os.path.dirname(HERE), "include")
    deps += [os.path.join(inc, h) for h in os.listdir(inc) if h.endswith(".h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=True, incremental=False):
    """Compile every HIP source and link the shared library.  Always recompiles, unless
    `incremental` is set: then objects kept in _lib/obj are reused when no source or header
    is newer (a developer convenience; the default build starts from scratch)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj") if incremental else OUT_DIR
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    todo = [(s, o) for s, o in zip(SOURCES, objs) if not incremental or _stale(s, o)]
    if todo:
        with ThreadPoolExecutor(max_workers=min(int(os.environ.get("MAX_JOBS", "8")), len(todo))) as ex:
            list(ex.map(lambda p: _compile(*p), todo))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if not incremental:
        for o in objs:
            os.remove(o)
    if verbose:
        print("built", LIB)
    build_flowio(verbose)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
