"""Flow files of the time loop (``flow_%010d.mat``, MAT 7.3) through libgatt_flowio.so.

The reference saves and loads them with MATLAB's MAT-file API
(``source/Fortran/readWriteMat.F90``, ``exportFlow.F90``, ``importFlow.F90``);
this module calls the library's HDF5 implementation of the same layout
(``include/gatt_flow_io.h``): variables ``t``, ``U``, ``V``, ``W``, ``R``, ``E``,
NaN inside walls on export, walls recognised from NaN on import.
"""
import ctypes as C
import os

import numpy as np

from .abi import DP, field_ptrs

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgatt_flowio.so")
EXPORTS = ["gatt_flow_last_error", "gatt_flow_write", "gatt_flow_read"]

_lib = None


class FlowFileError(RuntimeError):
    pass


def load_library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlowFileError("libgatt_flowio.so is not built (gatt_dns_amd.build.build_flowio)")
        L = C.CDLL(LIB_PATH)
        L.gatt_flow_last_error.restype = C.c_char_p
        L.gatt_flow_write.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.POINTER(DP),
                                      C.c_void_p]
        L.gatt_flow_read.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                     C.POINTER(DP), C.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise FlowFileError("flow file error %d: %s" % (rc, _lib.gatt_flow_last_error().decode()))


def flow_file_name(step):
    """``flow_%010d.mat``, or ``flow_diverged.mat`` for a negative step (readWriteMat.F90:172-179)."""
    return "flow_diverged.mat" if step < 0 else "flow_%010d.mat" % step


def write_flow(path, q, t, shape, flow_region=None):
    """Write state ``q`` (dict U..E of arrays of ``shape``) at time ``t``; points outside
    ``flow_region`` (True at flow points) are stored as NaN (exportFlow.F90)."""
    L = load_library()
    nx, ny, nz = shape
    arrs = [np.asfortranarray(np.asarray(q[k], dtype=np.float64).reshape(shape, order="F")) for k in "UVWRE"]
    wall = None
    if flow_region is not None:
        wall = np.asfortranarray(~np.asarray(flow_region, dtype=bool).reshape(shape, order="F")).astype(np.uint8)
    _check(L.gatt_flow_write(os.fsencode(path), nx, ny, nz, float(t), field_ptrs(arrs),
                             None if wall is None else wall.ctypes.data_as(C.c_void_p)))


def read_flow(path, shape):
    """Read a flow file: returns (state dict U..E, t, inside_wall bool array) with the
    importFlow.F90 replacements inside walls (U = V = W = 0, R = E = 1)."""
    L = load_library()
    nx, ny, nz = shape
    arrs = [np.zeros(shape, order="F") for _ in range(5)]
    wall = np.zeros(shape, dtype=np.uint8, order="F")
    t = C.c_double(0.0)
    _check(L.gatt_flow_read(os.fsencode(path), nx, ny, nz, C.byref(t), field_ptrs(arrs),
                            wall.ctypes.data_as(C.c_void_p)))
    return dict(zip("UVWRE", arrs)), t.value, wall.astype(bool)
