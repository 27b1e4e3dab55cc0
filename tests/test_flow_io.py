"""Flow files (flow_%010d.mat, MAT 7.3) of libgatt_flowio.so.

* Round trip: write then read gives the state bit for bit, the time, and the
  wall mask from the NaN convention (exportFlow.F90 writes NaN inside walls;
  importFlow.F90:24-40 reads them back as U = V = W = 0, R = E = 1).
* Layout, checked independently through libhdf5 (no MATLAB here): the file
  starts with the 128-byte MAT header ("MATLAB 7.3 MAT-file", version 0x0200,
  "IM") in a 512-byte userblock, the HDF5 signature sits at offset 512, every
  variable is a dataset with the attribute MATLAB_class = "double", 3-D arrays
  are stored with dimensions (nz, ny, nx) and 2-D ones as (ny, nx).
* The SFD mean (meanflowSFD.mat: U..E without t), the damping field (SFD.mat:
  SFD_X) and the flow with its rates (t, U..E, Udot..Edot) have the same layout,
  checked the same way, and read back bit for bit.
* log.txt: the header is runDNS.m:58-66's text; every field of a line is
  compared with what the image's Fortran compiler (amdflang) prints for the
  reference's own edit descriptors (main.F90:474, 499) on the same values,
  over random magnitudes, signs, zeros, NaN, Inf and fields too narrow for
  their value.
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
    _check_layout(path, [("t", [1, 1])] + [(k, dims) for k in "UVWRE"])


def _check_layout(path, variables):
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
    for name, want in variables:
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


@pytest.mark.parametrize("shape,dims", [((7, 5, 3), [3, 5, 7]), ((9, 6, 1), [6, 9])])
def test_mean_flow_sfd_and_rates(lib, tmp_path, shape, dims):
    q, qd = _state(shape, 5), _state(shape, 6)
    fr = np.ones(shape, dtype=bool)
    fr[-1, :, :] = False
    # meanflowSFD.mat: no t, NaN in walls; read back with the flow's mask or from the NaN
    path = tmp_path / "meanflowSFD.mat"
    flowio.write_mean_flow(path, q, shape, flow_region=fr)
    _check_layout(path, [(k, dims) for k in "UVWRE"])
    for mask in (~fr, None):
        got = flowio.read_mean_flow(path, shape, inside_wall=mask)
        for k, v in zip("UVWRE", (0.0, 0.0, 0.0, 1.0, 1.0)):
            np.testing.assert_array_equal(got[k][fr], q[k][fr])
            assert np.all(got[k][~fr] == v)
    with pytest.raises(flowio.FlowFileError):        # a mean file is not a flow file (no t)
        flowio.read_flow(path, shape)
    # SFD.mat
    x = np.asfortranarray(np.random.Generator(np.random.PCG64(7)).random(shape))
    path = tmp_path / "SFD.mat"
    flowio.write_sfd(path, x if shape[2] > 1 else x[:, :, 0])
    _check_layout(path, [("SFD_X", dims)])
    np.testing.assert_array_equal(flowio.read_sfd(path, shape), x)
    with pytest.raises(flowio.FlowFileError):
        flowio.read_sfd(path, (shape[0] + 1,) + shape[1:])
    # the flow with its rates: the flow part reads back as a flow file
    path = tmp_path / "flow_0000000004.mat"
    flowio.write_flow_deriv(path, q, {k + "dot": qd[k] for k in "UVWRE"}, 2.5, shape, flow_region=fr)
    _check_layout(path, [("t", [1, 1])] + [(k, dims) for k in "UVWRE"] + [(k + "dot", dims) for k in "UVWRE"])
    got, t, wall = flowio.read_flow(path, shape)
    assert t == 2.5
    np.testing.assert_array_equal(wall, ~fr)
    np.testing.assert_array_equal(got["E"][fr], q["E"][fr])


LOG_HEADER = ("Save number\tIteration\tSimulation time\tdt       \tCFL      \tU change\tV change\tW change"
              "\tR change\tE change")


def test_log_header(lib):
    assert flowio.log_header(0) == LOG_HEADER + "\n"
    two = flowio.log_header(2)
    cols = two.rstrip("\n").split("\t")
    assert cols[:10] == LOG_HEADER.split("\t")
    assert cols[10:] == [v + str(i) + " " * 12 for i in (1, 2) for v in "UVWRE"]


def test_log_tracked_values(lib):
    """main.F90:490-501: with a tracked norm R - 1 and E/norm - 1; exact zeros are "0"."""
    line = flowio.log_line(0, 7, 0.25, 1e-4, 0.9, [1e-3] * 5, [[0.0, 0.5, 0.0, 1.0, 2.0]], tracked_norm=4.0)
    f = line.rstrip(" \n").split("\t")
    assert f[10:] == ["0", " 5.00000000E-01", "0", "0", "-5.00000000E-01"]
    line = flowio.log_line(5, 7, 0.25, 1e-4, 0.9, [1e-3] * 5, [[0.0, 0.5, 0.0, 1.0, 2.0]])
    f = line.rstrip(" \n").split("\t")
    assert f[0] == "         5" and f[10:] == ["0", " 5.00000000E-01", "0", " 1.00000000E+00", " 2.00000000E+00"]
    assert line.endswith(" \n")


LOG_FORTRAN = """
program logfmt
  implicit none
  integer :: n, i, nsave, tstep
  real*8 :: v(10)
  open(1, file='in.bin', access='stream', form='unformatted', status='old')
  open(2, file='out.txt', status='replace')
  read(1) n
  do i = 1, n
    read(1) nsave, tstep, v
    write(2,'(I10,A1,I10,A1,ES13.8E1,A1,ES10.5E1,A1,F8.5,5(A1,ES10.4E2))',advance='no') nsave, char(9), tstep, &
      & char(9), v(1), char(9), v(2), char(9), v(3), char(9), v(4), char(9), v(5), char(9), v(6), char(9), v(7), &
      & char(9), v(8)
    write(2,'(A1,ES15.8E2,A1,ES15.8E2)',advance='no') char(9), v(9), char(9), v(10)
    write(2,'(A)') ''
  enddo
  close(2)
end program
"""


def _log_values(rng, count):
    """Rows of (nsave, tstep, t, dt, cfl, 5 changes, 2 tracked values) over many magnitudes."""
    rows = []
    specials = [0.0, -0.0, np.nan, np.inf, -np.inf, 9.99995, 9.999949999, 0.5, 123.5, 1e-10, 9.5e9, 1e100, -1e-100]
    for i in range(count):
        mag = 10.0 ** rng.uniform(-14, 14, 10)
        sgn = np.where(rng.random(10) < 0.3, -1.0, 1.0)
        v = mag * sgn
        if i % 4 == 0:
            v[rng.integers(0, 10, 3)] = rng.choice(specials, 3)
        if i % 3 == 0:
            v[2:5] = np.abs(v[2:5]) * np.array([1e-9, 1e-4, 1e-9])    # cfl in the F8.5 range
        nsave = int(rng.integers(0, 2 ** 31)) * (1 if i % 5 else -1)          # I10 overflows at -10^9
        rows.append((nsave, int(rng.integers(0, 2 ** 31)), v))
    return rows


def test_log_line_matches_fortran_edit_descriptors(lib, tmp_path):
    import struct
    import subprocess
    from gatt_dns_amd.fortran.build_driver import FC
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    rows = _log_values(np.random.Generator(np.random.PCG64(11)), 400)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(struct.pack("<i", len(rows)))
        for nsave, tstep, v in rows:
            f.write(struct.pack("<ii", nsave, tstep) + np.asarray(v, "<f8").tobytes())
    (tmp_path / "t.f90").write_text(LOG_FORTRAN)
    r = subprocess.run([FC, "-O0", "t.f90", "-o", "t"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run(["./t"], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stderr
    want = (tmp_path / "out.txt").read_text().split("\n")
    for k, (nsave, tstep, v) in enumerate(rows):
        tracked = [[v[8], v[9], 0.0, 0.0, 0.0]]
        line = flowio.log_line(nsave, tstep, v[0], v[1], v[2], v[3:8], tracked)
        got = line[:-2].split("\t")                                # drop " \n"
        exp = want[k].split("\t")
        # the tracked zeros are "0" by the reference's branch, not an edit descriptor
        tr = [exp[10] if v[8] != 0 else "0", exp[11] if v[9] != 0 else "0"]
        assert got[:10] == exp[:10], (k, v)
        assert got[10:12] == tr, (k, v)


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
  real(c_double) :: t2, mc(5), tr(5)
  integer(c_int) :: rc, i
  character(kind=c_char, len=256) :: line
  character(len=110) :: ref
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
  rc = gatt_meanflow_write('meanflowSFD.mat'//c_null_char, nx, ny, nz, q, c_loc(wall))
  if (rc /= 0) stop 7
  rc = gatt_meanflow_read('meanflowSFD.mat'//c_null_char, nx, ny, nz, q2, c_loc(wall))
  if (rc /= 0) stop 8
  if (any(V2(2:,:,:) /= V(2:,:,:)) .or. any(E2(1,:,:) /= 1d0)) stop 9
  rc = gatt_sfd_write('SFD.mat'//c_null_char, nx, ny, nz, R)
  if (rc /= 0) stop 10
  rc = gatt_sfd_read('SFD.mat'//c_null_char, nx, ny, nz, R2)
  if (rc /= 0 .or. any(R2 /= R)) stop 11
  mc = (/ 1d-3, 2d-3, 0d0, 4d-3, 5d-3 /)
  tr = (/ 1d0, 2d0, 0d0, 1d0, 1d0 /)
  rc = gatt_log_line(0, 10, 0.5d0, 1d-4, 0.25d0, mc, 1, tr, 0d0, line, int(len(line), c_int64_t))
  if (rc /= 178) stop 12
  write(ref,'(I10,A1,I10,A1,ES13.8E1,A1,ES10.5E1,A1,F8.5,5(A1,ES10.4E2))') 0, char(9), 10, char(9), 0.5d0, &
    & char(9), 1d-4, char(9), 0.25d0, char(9), mc(1), char(9), mc(2), char(9), mc(3), char(9), mc(4), char(9), mc(5)
  do i = 1, 110
    if (line(i:i) /= ref(i:i)) stop 13
  end do
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


def _read_dataset(h, f, name, complex_=False):
    d = h.H5Dopen2(f, name.encode(), 0)
    assert d >= 0, name
    sp = h.H5Dget_space(d)
    nd = h.H5Sget_simple_extent_ndims(sp)
    dims = (C.c_uint64 * nd)()
    h.H5Sget_simple_extent_dims(sp, dims, None)
    n = int(np.prod(list(dims)))
    native = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    h.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    if complex_:
        h.H5Tcreate.restype = C.c_int64
        h.H5Tcreate.argtypes = [C.c_int, C.c_size_t]
        h.H5Tinsert.argtypes = [C.c_int64, C.c_char_p, C.c_size_t, C.c_int64]
        t = h.H5Tcreate(6, 16)                                           # H5T_COMPOUND
        h.H5Tinsert(t, b"real", 0, native)
        h.H5Tinsert(t, b"imag", 8, native)
        buf = np.zeros(2 * n)
        assert h.H5Dread(d, t, 0, 0, 0, buf.ctypes.data) >= 0
        h.H5Tclose(t)
        data = buf[0::2] + 1j * buf[1::2]
    else:
        buf = np.zeros(n)
        assert h.H5Dread(d, native, 0, 0, 0, buf.ctypes.data) >= 0
        data = buf
    h.H5Sclose(sp)
    h.H5Dclose(d)
    return list(dims), data


def test_mat_write_real_and_complex(lib, tmp_path):
    """gatt_mat_write: MATLAB sizes reversed into HDF5 dimensions, column-major data, complex
    values as the compound {real, imag}, 1-D arrays as columns and scalars as 1x1."""
    rng = np.random.Generator(np.random.PCG64(8))
    a = np.asfortranarray(rng.standard_normal((4, 3, 1, 2)))
    z = np.asfortranarray(rng.standard_normal((5, 2)) + 1j * rng.standard_normal((5, 2)))
    v = rng.standard_normal(6)
    path = tmp_path / "m.mat"
    flowio.write_mat(path, {"A": a, "Z": z, "v": v, "s": 3.0})
    raw = open(path, "rb").read()
    assert raw[:19] == b"MATLAB 7.3 MAT-file" and raw[512:516] == b"\x89HDF"
    h = _h5()
    f = h.H5Fopen(os.fsencode(path), 0, 0)
    dims, data = _read_dataset(h, f, "A")
    assert dims == [2, 1, 3, 4]
    np.testing.assert_array_equal(data, a.reshape(-1, order="F"))
    dims, data = _read_dataset(h, f, "Z", complex_=True)
    assert dims == [2, 5]
    np.testing.assert_array_equal(data, z.reshape(-1, order="F"))
    dims, data = _read_dataset(h, f, "v")
    assert dims == [1, 6]
    np.testing.assert_array_equal(data, v)
    dims, data = _read_dataset(h, f, "s")
    assert dims == [1, 1] and data[0] == 3.0
    h.H5Fclose(f)


def test_modes_file(lib, tmp_path):
    """vec2modes: each mode divided by its largest modulus, NaN off the region, W NaN in 2-D,
    modes of eigenvalues with negative imaginary part left NaN."""
    from types import SimpleNamespace
    from gatt_dns_amd import lst
    case = SimpleNamespace(nx=6, ny=5, nz=1, X=np.linspace(0, 1, 6), Y=np.linspace(0, 2, 5), Z=np.zeros(1))
    region = np.array([1, 2, 3, 7, 8, 9, 20, 21])
    rng = np.random.Generator(np.random.PCG64(9))
    modes = []
    for _ in range(3):
        re = {f: rng.standard_normal((6, 5, 1)) for f in "UVWRE"}
        im = {f: rng.standard_normal((6, 5, 1)) for f in "UVWRE"}
        modes.append((re, im))
    out = {"modes": modes, "lambda": np.array([0.1 + 0.5j, 0.1 - 0.5j, -0.2 + 0.0j, -0.3 + 0.1j])}
    fields = lst.modes_fields(case, region, out)
    assert fields["U"].shape == (6, 5, 1, 3)
    off = np.ones(30, bool)
    off[region] = False
    for i in (0, 2):
        vals = np.concatenate([fields[f][..., i].reshape(-1, order="F")[region] for f in "UVRE"])
        assert np.max(np.abs(vals)) == pytest.approx(1.0, rel=1e-15)
        want = (modes[i][0]["V"] + 1j * modes[i][1]["V"]).reshape(-1, order="F")[region]
        ratio = fields["V"][..., i].reshape(-1, order="F")[region] / want
        np.testing.assert_allclose(ratio, ratio[0], rtol=1e-14)
        assert np.all(np.isnan(fields["U"][..., i].reshape(-1, order="F")[off]))
        assert np.all(np.isnan(fields["W"][..., i]))
    assert np.all(np.isnan(fields["U"][..., 1]))
    path = tmp_path / "modes-case-M4.mat"
    lst.write_modes(path, case, region, out)
    h = _h5()
    f = h.H5Fopen(os.fsencode(path), 0, 0)
    dims, data = _read_dataset(h, f, "U", complex_=True)
    assert dims == [3, 1, 5, 6]
    np.testing.assert_array_equal(np.isnan(data), np.isnan(fields["U"].reshape(-1, order="F")))
    dims, data = _read_dataset(h, f, "lambda", complex_=True)
    assert dims == [1, 4]
    np.testing.assert_array_equal(data, out["lambda"])
    h.H5Fclose(f)


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("shape", [(7, 5, 3), (9, 6, 1)])
def test_sfd_read_mat_v5(lib, tmp_path, compress, shape):
    """readSFD of the SFD.mat preprocessing.m:104 writes in MATLAB's default format (v7: level 5
    MAT file, zlib-compressed variables), checked against scipy's independent writer of that
    format; another variable stored first, a size mismatch and a missing SFD_X are errors or skipped
    as they should be.  (MATLAB may store integer-valued doubles in a narrower type; scipy does not,
    so the narrow-type path is restated from the format description only.)"""
    import scipy.io
    x = np.asfortranarray(np.random.Generator(np.random.PCG64(12)).random(shape))
    path = tmp_path / "SFD.mat"
    arr = x if shape[2] > 1 else x[:, :, 0]
    scipy.io.savemat(str(path), {"other": np.arange(5.0), "SFD_X": arr}, do_compression=compress, format="5")
    np.testing.assert_array_equal(flowio.read_sfd(path, shape), x)
    with pytest.raises(flowio.FlowFileError):
        flowio.read_sfd(path, (shape[0] + 1,) + shape[1:])
    scipy.io.savemat(str(path), {"other": np.arange(5.0)}, do_compression=compress, format="5")
    with pytest.raises(flowio.FlowFileError, match="no SFD_X"):
        flowio.read_sfd(path, shape)


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("shape", [(7, 5, 3), (9, 6, 1)])
def test_flow_read_mat_v5(lib, tmp_path, compress, shape):
    """The averaged flow files of runDNSautomean.m:80 are saved in MATLAB's default format (v7,
    level 5); gatt_flow_read takes them as it takes MAT 7.3 files: t and U..E (W may be absent),
    NaN inside walls replaced as importFlow.F90 does.  Written with scipy's independent writer."""
    import scipy.io
    rng = np.random.Generator(np.random.PCG64(5))
    q = {k: np.asfortranarray(rng.random(shape)) for k in "UVWRE"}
    q["U"][0, 0, 0] = np.nan
    squeeze = (lambda a: a) if shape[2] > 1 else (lambda a: a[:, :, 0])
    path = tmp_path / "flow_0000000009.mat"
    data = {"t": np.array([[2.5]]), **{k: squeeze(v) for k, v in q.items()}}
    if shape[2] == 1:
        del data["W"]
    scipy.io.savemat(str(path), data, do_compression=compress, format="5")
    got, t, wall = flowio.read_flow(path, shape)
    assert t == 2.5
    assert wall[0, 0, 0] and wall.sum() == 1
    for k in "UVWRE":
        want = q[k].copy() if (k != "W" or shape[2] > 1) else np.zeros(shape)
        want[0, 0, 0] = {"U": 0.0, "V": 0.0, "W": 0.0, "R": 1.0, "E": 1.0}[k]
        np.testing.assert_array_equal(got[k], want)
    scipy.io.savemat(str(path), {"t": np.array([[1.0]]), "U": squeeze(q["U"])}, format="5")
    with pytest.raises(flowio.FlowFileError):
        flowio.read_flow(path, shape)


def test_sfd_read_truncated_v5(lib, tmp_path):
    """A level 5 file cut inside an element's padding or data is an error, not a read past the buffer."""
    import scipy.io
    path = tmp_path / "SFD.mat"
    scipy.io.savemat(str(path), {"SFD_X": np.arange(21.0).reshape(7, 3)}, format="5")
    raw = path.read_bytes()
    for cut in (len(raw) - 3, len(raw) - 20, 140, 131):
        path.write_bytes(raw[:cut])
        with pytest.raises(flowio.FlowFileError):
            flowio.read_sfd(path, (7, 3, 1))
