"""ctypes mirror of ``include/gatt_dns.h`` and the descriptor builder.

:func:`build_desc` turns a :class:`~gatt_dns_amd.prep.cases.Case` into a
``gatt_problem_desc`` whose pointers reference numpy arrays kept alive by the
returned :class:`Desc` object.  All arrays are laid out as the reference's
Fortran arrays (column-major).
"""
import ctypes as C

import numpy as np

I32 = C.c_int32
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)


class LineOp(C.Structure):
    _fields_ = [("n", I32), ("ntypes", I32), ("nR", I32), ("periodic", I32),
                ("A", DP), ("B", DP), ("C", DP), ("D", DP), ("R", DP)]


class BoxList(C.Structure):
    _fields_ = [("count", I32), ("box", IP), ("value", DP), ("dir", IP)]


class Disturbance(C.Structure):
    _fields_ = [("kind", I32), ("box", I32 * 6), ("x", DP), ("par", C.c_double * 4), ("var", I32)]


class Forcing(C.Structure):
    _fields_ = [("kind", I32), ("box", I32 * 6), ("par", C.c_double * 4), ("var", I32 * 2)]


class BCDesc(C.Structure):
    _fields_ = [("dir", BoxList * 5), ("neu", BoxList * 5), ("sec", BoxList * 5),
                ("nCorners", I32), ("cL", IP), ("cD", IP), ("cAdiabatic", IP),
                ("neumannLength", I32), ("neumann2Length", I32),
                ("neumannCoeffs", DP), ("neumann2Coeffs", DP),
                ("nDisturb", I32), ("disturb", C.POINTER(Disturbance))]


class ProblemDesc(C.Structure):
    _fields_ = [("nx", I32), ("ny", I32), ("nz", I32),
                ("Re", C.c_double), ("Ma", C.c_double), ("Pr", C.c_double), ("gamma", C.c_double),
                ("T0", C.c_double),
                ("deriv", LineOp * 3), ("filter", LineOp * 3), ("types", IP * 3),
                ("filterDir", I32 * 3), ("filterCharTime", C.c_double),
                ("bc", BCDesc), ("nForcing", I32), ("forcing", C.POINTER(Forcing)),
                ("sfdType", I32), ("sfdDelta", C.c_double), ("sfdXval", C.c_double), ("sfdX", DP),
                ("dxmin", C.c_double * 3), ("timeStepping", I32)]


GATT_U, GATT_V, GATT_W, GATT_R, GATT_E = range(5)
GATT_DERIV, GATT_FILTER = 0, 1
BC_SLOTS = "UVWPE"
DIST_KIND = {"lidDrivenMovement": 1, "holdInlet": 2, "periodicSine": 3, "packet_2d": 4}
FORCING_KIND = {"turbulenceGenerator": 1, "pressureGradient": 2}
FIELD_INDEX = {"U": GATT_U, "V": GATT_V, "W": GATT_W, "R": GATT_R, "E": GATT_E}


class Desc:
    """Owns the numpy arrays a :class:`ProblemDesc` points to."""

    def __init__(self):
        self.keep = []
        self.desc = ProblemDesc()

    def dptr(self, a):
        if a is None:
            return DP()
        arr = np.asfortranarray(np.asarray(a, dtype=np.float64))
        self.keep.append(arr)
        return arr.ctypes.data_as(DP)

    def iptr(self, a):
        if a is None:
            return IP()
        arr = np.asfortranarray(np.asarray(a, dtype=np.int32))
        self.keep.append(arr)
        return arr.ctypes.data_as(IP)

    @property
    def ptr(self):
        return C.byref(self.desc)


def _lineop(d, op):
    lo = LineOp()
    if op is None:
        return lo
    lo.n, lo.ntypes, lo.nR, lo.periodic = op.n, op.ntypes, op.nR, int(op.periodic)
    lo.A, lo.B, lo.C = d.dptr(op.A), d.dptr(op.B), d.dptr(op.C)
    lo.D = d.dptr(op.D) if op.periodic else DP()
    lo.R = d.dptr(op.R)
    return lo


def build_desc(case, sfd_override=None):
    """Build the ``gatt_problem_desc`` of a case."""
    d = Desc()
    p = d.desc
    p.nx, p.ny, p.nz = case.nx, case.ny, case.nz
    p.Re, p.Ma, p.Pr, p.gamma, p.T0 = case.Re, case.Ma, case.Pr, case.gamma, case.T0
    for k in range(3):
        p.deriv[k] = _lineop(d, case.deriv[k])
        p.filter[k] = _lineop(d, case.filt[k])
        p.types[k] = d.iptr(case.types[k])
        p.filterDir[k] = int(case.filter_dirs[k])
    p.filterCharTime = case.filter_char_time
    p.timeStepping = int(getattr(case, "time_stepping", 1))
    bi = case.bi
    for s, v in enumerate(BC_SLOTS):
        boxes, vals = bi.d[v]
        p.bc.dir[s] = BoxList(len(boxes), d.iptr(boxes), d.dptr(vals), IP())
        boxes, dirs = bi.n[v]
        p.bc.neu[s] = BoxList(len(boxes), d.iptr(boxes), DP(), d.iptr(dirs))
        boxes, dirs = bi.s[v]
        p.bc.sec[s] = BoxList(len(boxes), d.iptr(boxes), DP(), d.iptr(dirs))
    p.bc.nCorners = len(bi.cL)
    p.bc.cL, p.bc.cD, p.bc.cAdiabatic = d.iptr(bi.cL), d.iptr(bi.cD), d.iptr(bi.cAdiabatic)
    p.bc.neumannLength, p.bc.neumann2Length = len(bi.neumann), len(bi.neumann2)
    p.bc.neumannCoeffs, p.bc.neumann2Coeffs = d.dptr(bi.neumann), d.dptr(bi.neumann2)
    nd = len(case.disturb)
    darr = (Disturbance * max(nd, 1))()
    for i, ds in enumerate(case.disturb):
        darr[i].kind = DIST_KIND[ds.kind]
        for c in range(6):
            darr[i].box[c] = ds.box[c]
        darr[i].x = d.dptr(ds.X)
        for c, v in enumerate(ds.par[:4]):
            darr[i].par[c] = v
        if ds.var:
            darr[i].var = FIELD_INDEX[ds.var[0].upper()]
    d.keep.append(darr)
    p.bc.nDisturb = nd
    p.bc.disturb = C.cast(darr, C.POINTER(Disturbance))
    nf = len(case.forcing)
    farr = (Forcing * max(nf, 1))()
    for i, fc in enumerate(case.forcing):
        farr[i].kind = FORCING_KIND[fc.kind]
        for c in range(6):
            farr[i].box[c] = fc.box[c]
        for c, v in enumerate(fc.par[:4]):
            farr[i].par[c] = v
        for c, v in enumerate(fc.var[:2]):
            farr[i].var[c] = FIELD_INDEX[v.upper()]
    d.keep.append(farr)
    p.nForcing = nf
    p.forcing = C.cast(farr, C.POINTER(Forcing))
    p.sfdType, p.sfdDelta, p.sfdXval = case.sfd_type, case.sfd_delta, case.sfd_x_val
    p.sfdX = d.dptr(case.sfd_x) if case.sfd_x is not None else DP()
    for k in range(3):
        p.dxmin[k] = case.dxmin[k]
    return d


def field_ptrs(arrs, ctype=DP):
    """Array of 5 pointers (U,V,W,R,E) to Fortran-ordered float64 arrays (None -> NULL).

    The returned object holds references to the arrays, so temporaries stay
    alive for as long as the pointer array is in use.
    """
    arrs = list(arrs)
    out = (ctype * 5)()
    out._keep = arrs
    for i, a in enumerate(arrs):
        if a is None:
            out[i] = ctype()
        else:
            assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
            out[i] = a.ctypes.data_as(ctype)
    return out
