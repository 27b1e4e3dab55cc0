"""Flow files (flow_%010d.mat, MAT 7.3) of libgatt_flowio.so.

* Round trip: write then read gives the state bit for bit, the time, and the
  wall mask from the NaN convention (exportFlow.F90 writes NaN inside walls;
  importFlow.F90:24-40 reads them back as U = V = W = 0, R = E = 1).
* Layout, checked independently through libhdf5 (no MATLAB here): the file
  starts with the 128-byte MAT header ("MATLAB 7.3 MAT-file", version 0x0200,
  "IM") in a 512-byte userblock, the HDF5 signature sits at offset 512, every
  variable is a dataset with the attribute MATLAB_class = "double", 3-D arrays
  are stored with dimensions (nz, ny, nx) and 2-D ones as (ny, nx).
* Compatibility with files written by MATLAB itself is not checked (parity
  unpinned): the reference ships no flow files and MATLAB is not available.
"""
import ctypes as C
import os

import numpy as np
import pytest

from gatt_dns_amd import flowio

HDF5 = "/opt/conda/lib/libhdf5.so"


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(flowio.LIB_PATH):
        from gatt_dns_amd import build
        if build.build_flowio(verbose=False) is None:
            pytest.skip("no HDF5 headers")
    return flowio.load_library()


def _state(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return {k: np.asfortranarray(rng.standard_normal(shape)) for k in "UVWRE"}


@pytest.mark.parametrize("shape", [(7, 5, 3), (9, 6, 1)])
def test_round_trip_and_walls(lib, tmp_path, shape):
    q = _state(shape, 3)
    fr = np.ones(shape, dtype=bool)
    fr[0, :, :] = False
    fr[:, -1, :] = False
    path = tmp_path / flowio.flow_file_name(12)
    assert path.name == "flow_0000000012.mat"
    flowio.write_flow(path, q, 1.25, shape, flow_region=fr)
    got, t, wall = flowio.read_flow(path, shape)
    assert t == 1.25
    np.testing.assert_array_equal(wall, ~fr)
    for k in "UVWRE":
        np.testing.assert_array_equal(got[k][fr], q[k][fr])
    for k, v in zip("UVWRE", (0.0, 0.0, 0.0, 1.0, 1.0)):
        assert np.all(got[k][~fr] == v)
    # without a region every value travels unchanged
    flowio.write_flow(path, q, -3.0, shape)
    got, t, wall = flowio.read_flow(path, shape)
    assert t == -3.0 and not wall.any()
    for k in "UVWRE":
        np.testing.assert_array_equal(got[k], q[k])


def test_size_mismatch_is_an_error(lib, tmp_path):
    path = tmp_path / "flow_0000000000.mat"
    flowio.write_flow(path, _state((4, 3, 2), 1), 0.0, (4, 3, 2))
    with pytest.raises(flowio.FlowFileError):
        flowio.read_flow(path, (4, 3, 3))
    with pytest.raises(flowio.FlowFileError):
        flowio.read_flow(tmp_path / "missing.mat", (4, 3, 2))


def _h5():
    if not os.path.exists(HDF5):
        pytest.skip("libhdf5 not found")
    h = C.CDLL(HDF5)
    h.H5open()
    for fn, res in (("H5Fopen", C.c_int64), ("H5Dopen2", C.c_int64), ("H5Dget_space", C.c_int64),
                    ("H5Aopen", C.c_int64), ("H5Aget_type", C.c_int64)):
        getattr(h, fn).restype = res
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    h.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h.H5Dget_space.argtypes = [C.c_int64]
    h.H5Sget_simple_extent_ndims.argtypes = [C.c_int64]
    h.H5Sget_simple_extent_dims.argtypes = [C.c_int64, C.POINTER(C.c_uint64), C.c_void_p]
    h.H5Aopen.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h.H5Aget_type.argtypes = [C.c_int64]
    h.H5Tget_size.argtypes = [C.c_int64]
    h.H5Tget_size.restype = C.c_size_t
    h.H5Aread.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
    h.H5Pget_userblock.argtypes = [C.c_int64, C.POINTER(C.c_uint64)]
    h.H5Fget_create_plist.argtypes = [C.c_int64]
    h.H5Fget_create_plist.restype = C.c_int64
    for fn in ("H5Dclose", "H5Sclose", "H5Aclose", "H5Tclose", "H5Fclose", "H5Pclose"):
        getattr(h, fn).argtypes = [C.c_int64]
    return h


@pytest.mark.parametrize("shape,dims", [((7, 5, 3), [3, 5, 7]), ((9, 6, 1), [6, 9])])
def test_mat73_layout(lib, tmp_path, shape, dims):
    path = tmp_path / "flow_0000000001.mat"
    flowio.write_flow(path, _state(shape, 2), 0.5, shape)
    raw = open(path, "rb").read()
    assert raw[:19] == b"MATLAB 7.3 MAT-file"
    assert raw[124:128] == b"\x00\x02IM"
    assert raw[512:520] == b"\x89HDF\r\n\x1a\n"
    h = _h5()
    f = h.H5Fopen(os.fsencode(path), 0, 0)
    assert f >= 0
    cp = h.H5Fget_create_plist(f)
    ub = C.c_uint64(0)
    h.H5Pget_userblock(cp, C.byref(ub))
    h.H5Pclose(cp)
    assert ub.value == 512
    for name, want in [("t", [1, 1])] + [(k, dims) for k in "UVWRE"]:
        d = h.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        sp = h.H5Dget_space(d)
        nd = h.H5Sget_simple_extent_ndims(sp)
        buf = (C.c_uint64 * nd)()
        h.H5Sget_simple_extent_dims(sp, buf, None)
        assert list(buf) == want, (name, list(buf))
        a = h.H5Aopen(d, b"MATLAB_class", 0)
        assert a >= 0, name
        ty = h.H5Aget_type(a)
        n = h.H5Tget_size(ty)
        s = C.create_string_buffer(n + 1)
        h.H5Aread(a, ty, s)
        assert s.raw[:n] == b"double"
        for x, fn in ((ty, "H5Tclose"), (a, "H5Aclose"), (sp, "H5Sclose"), (d, "H5Dclose")):
            getattr(h, fn)(x)
    h.H5Fclose(f)


FORTRAN_PROGRAM = """
program t
  use iso_c_binding
  use gatt_hip
  implicit none
  integer, parameter :: nx = 5, ny = 4, nz = 3
  real(c_double), target :: U(nx,ny,nz), V(nx,ny,nz), W(nx,ny,nz), R(nx,ny,nz), E(nx,ny,nz)
  real(c_double), target :: U2(nx,ny,nz), V2(nx,ny,nz), W2(nx,ny,nz), R2(nx,ny,nz), E2(nx,ny,nz)
  integer(c_int8_t), target :: wall(nx,ny,nz), wall2(nx,ny,nz)
  type(c_ptr) :: q(5), q2(5)
  real(c_double) :: t2
  integer(c_int) :: rc
  call random_number(U); call random_number(V); call random_number(W)
  call random_number(R); call random_number(E)
  wall = 0_c_int8_t
  wall(1, :, :) = 1_c_int8_t
  q = (/ c_loc(U), c_loc(V), c_loc(W), c_loc(R), c_loc(E) /)
  q2 = (/ c_loc(U2), c_loc(V2), c_loc(W2), c_loc(R2), c_loc(E2) /)
  rc = gatt_flow_write('flow_0000000003.mat'//c_null_char, nx, ny, nz, 1.5d0, q, c_loc(wall))
  if (rc /= 0) stop 1
  rc = gatt_flow_read('flow_0000000003.mat'//c_null_char, nx, ny, nz, t2, q2, c_loc(wall2))
  if (rc /= 0) stop 2
  if (t2 /= 1.5d0) stop 3
  if (any(wall2 /= wall)) stop 4
  if (any(U2(2:,:,:) /= U(2:,:,:)) .or. any(E2(2:,:,:) /= E(2:,:,:))) stop 5
  if (any(U2(1,:,:) /= 0d0) .or. any(R2(1,:,:) /= 1d0)) stop 6
  print *, 'fortran flow io ok'
end program
"""


def test_fortran_interfaces(lib, tmp_path):
    """gatt_hip.f90's interfaces of gatt_flow_write / gatt_flow_read from a Fortran program (no GPU
    call: the program links the GPU library but only touches the flow files)."""
    import subprocess
    from gatt_dns_amd.fortran.build_driver import FC, LIB_DIR
    mod = os.path.join(LIB_DIR, "fortran_build", "gatt_hip.o")
    if not (os.path.exists(FC) and os.path.exists(mod) and os.path.exists(os.path.join(LIB_DIR, "libgatt_dns.so"))):
        pytest.skip("amdflang or the Fortran module / GPU library not built")
    src = tmp_path / "t.f90"
    src.write_text(FORTRAN_PROGRAM)
    exe = tmp_path / "t"
    r = subprocess.run([FC, "-O2", "-I", os.path.dirname(mod), str(src), mod, "-L", LIB_DIR, "-lgatt_flowio",
                        "-lgatt_dns", "-Wl,-rpath," + LIB_DIR, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0 and "fortran flow io ok" in r.stdout, r.stdout + r.stderr
