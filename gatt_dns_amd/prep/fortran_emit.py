"""Emit a case as Fortran include files for the Fortran driver.

The files follow the statement format of the reference's generated includes
(``writeFortranParameters.m``, ``writeFortranMatrices.m``,
``writeFortranBoundaries.m``): scalar declarations, ``allocate`` statements and
array values written as ``%.20fd0`` literals, column-major.  Long arrays are
assigned in chunks of a flat buffer (``gbuf``) and reshaped, which keeps every
line short for any Fortran compiler.  Derivative blocks are written as the
reference writes them (``[type j1 j2 k1 k2]`` rows, ``getMatrixTypeBlocks.m``).
``disturbInfo.F90`` has no reference counterpart: it fills the disturbance and
forcing lists of the C descriptor.
"""
import os

import numpy as np


def _lit(v):
    return "%.20fd0" % v


def _assign_real(lines, name, arr, chunk=6):
    flat = np.asarray(arr, dtype=np.float64).ravel(order="F")
    lines.append("    if (allocated(gbuf)) deallocate(gbuf)")
    lines.append("    allocate(gbuf(%d))" % max(len(flat), 1))
    for k in range(0, len(flat), chunk):
        part = flat[k:k + chunk]
        lines.append("    gbuf(%d:%d) = (/%s/)" % (k + 1, k + len(part), ",".join(_lit(v) for v in part)))
    lines.append("    %s = reshape(gbuf(1:%d), shape(%s))" % (name, len(flat), name))


def _assign_int(lines, name, arr, chunk=24):
    flat = np.asarray(arr, dtype=np.int64).ravel(order="F")
    lines.append("    if (allocated(ibuf)) deallocate(ibuf)")
    lines.append("    allocate(ibuf(%d))" % max(len(flat), 1))
    for k in range(0, len(flat), chunk):
        part = flat[k:k + chunk]
        lines.append("    ibuf(%d:%d) = (/%s/)" % (k + 1, k + len(part), ",".join(str(int(v)) for v in part)))
    if len(flat):
        lines.append("    %s = reshape(ibuf(1:%d), shape(%s))" % (name, len(flat), name))


def matrix_type_blocks(type_map, max_block=128):
    """Blocks ``[type j1 j2 k1 k2]`` of one rank (``getMatrixTypeBlocks.m:8-100``, p_row = p_col = 1)."""
    J, K = type_map.shape
    blocks = []
    for k in range(K):
        col = type_map[:, k]
        starts = [0] + [j for j in range(1, J) if col[j] != col[j - 1]]
        ends = [s - 1 for s in starts[1:]] + [J - 1]
        for a, b in zip(starts, ends):
            blocks.append([int(col[a]), a + 1, b + 1, k + 1, k + 1])
    i = 0
    while i < len(blocks) - 1:
        j = i + 1
        while j < len(blocks):
            if blocks[i][0:3] == blocks[j][0:3] and blocks[i][4] + 1 == blocks[j][3]:
                blocks[i][4] = blocks[j][4]
                del blocks[j]
            else:
                j += 1
        i += 1
    out = []
    for b in blocks:
        isz, jsz = b[2] - b[1] + 1, b[4] - b[3] + 1
        nsl = -(-isz * jsz // max_block)
        if nsl == 1:
            out.append(b)
            continue
        isl = -(-nsl // jsz)
        jsl = nsl // isl
        def slices(n, p):
            base, extra = n // p, n - (n // p) * p
            sizes = [base] * p
            for q in range(p - extra, p):
                sizes[q] += 1
            ends = np.cumsum(sizes)
            return [(int(e - s) + 1, int(e)) for s, e in zip(sizes, ends)]
        for (a1, a2) in slices(isz, isl):
            for (c1, c2) in slices(jsz, jsl):
                out.append([b[0], b[1] + a1 - 1, b[1] + a2 - 1, b[3] + c1 - 1, b[3] + c2 - 1])
    return np.array(out, dtype=np.int64).reshape(-1, 5)


def write_includes(case, outdir, dt=None):
    """Write parameters.F90, matrices.F90, boundaryInfo.F90 and disturbInfo.F90 into outdir."""
    os.makedirs(outdir, exist_ok=True)
    dt = case.dt if dt is None else dt
    p = []
    p.append("    integer :: nx = %d" % case.nx)
    p.append("    integer :: ny = %d" % case.ny)
    p.append("    integer :: nz = %d" % case.nz)
    for nm in ("Re", "Ma", "Pr", "T0", "gamma"):
        p.append("    real*8 :: %s = %s" % (nm, _lit(getattr(case, nm))))
    p.append("    real*8 :: dtmax = %s" % _lit(dt))
    p.append("    real*8 :: FilterCharTime = %s" % _lit(case.filter_char_time))
    p.append("    integer :: SFD = %d" % case.sfd_type)
    p.append("    real*8 :: SFD_Delta = %s" % _lit(case.sfd_delta))
    p.append("    real*8 :: SFD_X_val = %s" % _lit(case.sfd_x_val))
    p.append("    real*8,dimension(3) :: dxmin = (/%s,%s,%s/)" % tuple(_lit(v) for v in case.dxmin))
    p.append("    integer :: timeStepping = %d" % getattr(case, "time_stepping", 1))
    open(os.path.join(outdir, "parameters.F90"), "w").write("\n".join(p) + "\n")

    m = []
    ndir = 3 if case.is3d else 2
    for d, X in zip(range(ndir), "XYZ"):
        blocks = matrix_type_blocks(case.types[d])
        m.append("    nDerivBlocks%s = %d" % (X, len(blocks)))
        m.append("    allocate(derivBlocks%s(%d,5))" % (X, len(blocks)))
        _assign_int(m, "derivBlocks%s" % X, blocks)
    m.append("    filterX = %d" % case.filter_dirs[0])
    m.append("    filterY = %d" % case.filter_dirs[1])
    m.append("    filterZ = %d" % case.filter_dirs[2])
    for d, X in zip(range(ndir), "XYZ"):
        x = X.lower()
        der, fil = case.deriv[d], case.filt[d]
        m.append("    derivnRHS%s = %d" % (x, der.nR))
        m.append("    filternRHS%s = %d" % (x, fil.nR))
        m.append("    ntypes%s = %d" % (X, der.ntypes))
        for pre, op in (("derivs", der), ("filter", fil)):
            m.append("    allocate(%sA%s(%d,%d))" % (pre, X, op.A.shape[0], op.A.shape[1]))
            m.append("    allocate(%sB%s(%d,%d))" % (pre, X, op.B.shape[0], op.B.shape[1]))
            m.append("    allocate(%sC%s(%d,%d))" % (pre, X, op.C.shape[0], op.C.shape[1]))
            m.append("    allocate(%sR%s(%d,%d,%d))" % ((pre, X) + op.R.shape))
            for k in "ABCR":
                _assign_real(m, "%s%s%s" % (pre, k, X), getattr(op, k))
        m.append("    periodic%s = %d" % (X, int(der.periodic)))
        if der.periodic:
            for pre, op in (("derivs", der), ("filter", fil)):
                m.append("    allocate(%sD%s(%d,%d))" % (pre, X, op.D.shape[0], op.D.shape[1]))
                _assign_real(m, "%sD%s" % (pre, X), op.D)
    bi = case.bi
    m.append("    neumannLength = %d" % len(bi.neumann))
    m.append("    allocate(neumannCoeffs(%d))" % len(bi.neumann))
    _assign_real(m, "neumannCoeffs", bi.neumann)
    m.append("    neumann2Length = %d" % len(bi.neumann2))
    m.append("    allocate(neumann2Coeffs(%d))" % len(bi.neumann2))
    _assign_real(m, "neumann2Coeffs", bi.neumann2)
    open(os.path.join(outdir, "matrices.F90"), "w").write("\n".join(m) + "\n")

    b = []
    for v in "UVWEP":               # writeFortranBoundaries.m:11
        boxes, vals = bi.d[v]
        b.append("    n%sd = %d" % (v, len(boxes)))
        b.append("    allocate(i%sd(%d,6))" % (v, max(len(boxes), 1)))
        b.append("    allocate(v%sd(%d))" % (v, max(len(boxes), 1)))
        if len(boxes):
            _assign_int(b, "i%sd" % v, boxes)
            _assign_real(b, "v%sd" % v, vals)
    for kind, src in (("n", bi.n), ("s", bi.s)):
        for v in "UVWEP":
            boxes, dirs = src[v]
            b.append("    n%s%s = %d" % (v, kind, len(boxes)))
            b.append("    allocate(i%s%s(%d,6))" % (v, kind, max(len(boxes), 1)))
            b.append("    allocate(d%s%s(%d))" % (v, kind, max(len(boxes), 1)))
            if len(boxes):
                _assign_int(b, "i%s%s" % (v, kind), boxes)
                _assign_int(b, "d%s%s" % (v, kind), dirs)
    b.append("    cN = %d" % len(bi.cL))
    b.append("    allocate(cL(%d,6))" % max(len(bi.cL), 1))
    b.append("    allocate(cD(%d,3))" % max(len(bi.cL), 1))
    b.append("    allocate(cAdiabatic(%d))" % max(len(bi.cL), 1))
    if len(bi.cL):
        _assign_int(b, "cL", bi.cL)
        _assign_int(b, "cD", bi.cD)
        _assign_int(b, "cAdiabatic", bi.cAdiabatic)
    open(os.path.join(outdir, "boundaryInfo.F90"), "w").write("\n".join(b) + "\n")

    kinds = {"lidDrivenMovement": 1, "holdInlet": 2, "periodicSine": 3, "packet_2d": 4}
    fields = {"U": 0, "V": 1, "W": 2, "R": 3, "E": 4}
    t = ["    nDist = %d" % len(case.disturb), "    allocate(dist(max(nDist,1)))"]
    nx_tot = sum(len(ds.X) for ds in case.disturb)
    t.append("    allocate(distX(max(%d,1)))" % nx_tot)
    off = 0
    for i, ds in enumerate(case.disturb):
        t.append("    dist(%d)%%kind = %d" % (i + 1, kinds[ds.kind]))
        t.append("    dist(%d)%%box = (/%s/)" % (i + 1, ",".join(str(v) for v in ds.box)))
        for k, v in enumerate(ds.par[:4]):
            t.append("    dist(%d)%%par(%d) = %s" % (i + 1, k + 1, _lit(v)))
        if ds.var:
            t.append("    dist(%d)%%var = %d" % (i + 1, fields[ds.var[0].upper()]))
        _assign_real(t, "distX(%d:%d)" % (off + 1, off + len(ds.X)), ds.X)
        t.append("    distOff(%d) = %d" % (i + 1, off + 1))
        off += len(ds.X)
    t.append("    nForc = %d" % len(case.forcing))
    t.append("    allocate(forc(max(nForc,1)))")
    for i, fc in enumerate(case.forcing):
        t.append("    forc(%d)%%kind = %d" % (i + 1, {"turbulenceGenerator": 1, "pressureGradient": 2}[fc.kind]))
        for k, v in enumerate(fc.var[:2]):
            t.append("    forc(%d)%%var(%d) = %d" % (i + 1, k + 1, fields[v.upper()]))
        t.append("    forc(%d)%%box = (/%s/)" % (i + 1, ",".join(str(v) for v in fc.box)))
        for k, v in enumerate(fc.par[:4]):
            t.append("    forc(%d)%%par(%d) = %s" % (i + 1, k + 1, _lit(v)))
    open(os.path.join(outdir, "disturbInfo.F90"), "w").write("\n".join(t) + "\n")
