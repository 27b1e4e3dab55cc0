"""Compact-scheme matrices and their pre-factored line-solve coefficients.

Host-side restatement of the reference preprocessing that produces every
coefficient the time-advance path consumes:

  * ``findDerivativeRegions.m``   -> :func:`find_derivative_regions`
  * ``makeMatrices.m``            -> :func:`make_matrices` (with
    ``makeMatricesEachDirection``, ``addBufferToMatrix`` and ``applyMetric``)
  * ``prepareThomas.m``           -> :func:`prepare_thomas`
  * ``writeFortranMatrices.m:228-235`` (``%.20fd0`` literals) -> :func:`fortran_round`

Index conventions: line starts/ends and break points are kept 1-based exactly as
in the reference; dense matrices and numpy arrays are 0-based.
"""
from dataclasses import dataclass, field

import numpy as np

from .stencils import finite_difference_coefficients, spatial_filter_coefficients


def fortran_round(a):
    """Value a Fortran compiler reads back from ``sprintf('%.20fd0', a)``.

    The reference never passes binary coefficients to the solver: they travel as
    fixed-point text with 20 decimals (``writeFortranMatrices.m:228-235``,
    ``writeFortranParameters.m:11-18``), so every value is rounded to 1e-20
    absolute before use.  Both printing and parsing are correctly rounded.
    """
    arr = np.asarray(a, dtype=np.float64)
    flat = np.array([float("%.20f" % v) for v in arr.ravel()], dtype=np.float64)
    return flat.reshape(arr.shape)


def matlab_linspace(d1, d2, n):
    """MATLAB ``linspace``: ``d1 + (0:n-1).*(d2-d1)/(n-1)`` with the last point set to d2."""
    if n == 1:
        return np.array([float(d2)])
    k = np.arange(n, dtype=np.float64)
    y = d1 + k * (d2 - d1) / (n - 1)
    y[-1] = d2
    return y


@dataclass
class BufferSide:
    n: int = 0
    upwind: bool = False
    transition: float = None


@dataclass
class DirMesh:
    """One direction of the mesh (``mesh.x`` / ``mesh.y`` / ``mesh.z`` in the reference)."""
    X: np.ndarray
    periodic: bool = False
    fix_periodic_domain_size: bool = False
    xf: float = None                      # domain end, used by applyMetric for periodic meshes
    buffer_i: BufferSide = field(default_factory=BufferSide)
    buffer_f: BufferSide = field(default_factory=BufferSide)
    break_points: list = field(default_factory=list)   # rows [position, a1, a2, b1, b2] (1-based)

    @property
    def n(self):
        return len(self.X)


@dataclass
class NumMethods:
    spatial_derivs: str = "SL4"
    spatial_derivs_buffer: str = "EX4"
    filter_strength: float = 0.49
    filter_borders: object = "decentered"
    filter_borders_start: tuple = (True, True, True)   # filterBordersStartX/Y/Z
    filter_borders_end: tuple = (True, True, True)     # filterBordersEndX/Y/Z
    change_order: tuple = (True, True, True)           # changeOrderX/Y/Z
    metric_method: str = "SL4"
    neumann_order: int = 6
    neumann2_order: int = 2
    filter_directions: tuple = (1, 1, 1)
    filter_char_time: float = 0.0                      # spatialFilterTime


# ---------------------------------------------------------------------------
# findDerivativeRegions.m
# ---------------------------------------------------------------------------

def _regions_one_dir(lines, periodic, n, break_points):
    """Type map and wall starts/ends for one direction.

    ``lines`` has shape (L1, n, L2): ``lines[a, :, b]`` is the flow-region mask along
    one line.  Follows ``findDerivativeRegions.m:5-62`` (the x block; y and z are the
    same code on permuted arrays).
    """
    L1, _, L2 = lines.shape
    rows = lines.transpose(0, 2, 1).reshape(-1, n).astype(np.int8)
    C = np.unique(rows, axis=0)                       # unique(...,'rows','sorted')
    type_map = np.zeros((L1, L2), dtype=np.int64)
    # typeMap(a,b) = row of C equal to the line (1-based)
    lut = {C[r].tobytes(): r + 1 for r in range(C.shape[0])}
    for a in range(L1):
        for b in range(L2):
            type_map[a, b] = lut[lines[a, :, b].astype(np.int8).tobytes()]
    starts, ends = [], []
    for i in range(1, C.shape[0] + 1):
        row = C[i - 1]
        if np.any(row != 0):
            d = np.diff(row.astype(np.int64))
            up = list(np.nonzero(d == 1)[0] + 1)          # find(diff==1), 1-based
            dn = list(np.nonzero(d == -1)[0] + 1 + 1)     # find(diff==-1)+1
            if periodic:
                starts.append(sorted(set(up)))
                ends.append(sorted(set(dn)))
            else:
                starts.append(sorted(set([1] + up)))
                ends.append(sorted(set(dn + [n])))
        else:
            # findDerivativeRegions.m:38-39; an all-wall row sorts first, so i == 1 here
            mask_gt_i = type_map > i
            mask_gt_1 = type_map > 1
            type_map[type_map == i] = 0
            type_map[mask_gt_i] = type_map[mask_gt_1] - 1
    type_map[type_map == 0] = type_map.max()
    for bp in break_points:                           # findDerivativeRegions.m:45-62
        pos, a1, a2, b1, b2 = [int(v) for v in bp]
        types = np.unique(type_map[a1 - 1:a2, b1 - 1:b2])
        for t in types:
            type_n = int(type_map.max()) + 1
            while len(starts) < type_n:
                starts.append([])
                ends.append([])
            starts[type_n - 1] = list(starts[t - 1]) + [pos + 1]
            ends[type_n - 1] = list(ends[t - 1]) + [pos]
            sub = type_map[a1 - 1:a2, b1 - 1:b2]
            sub[sub == t] = type_n
            type_map[a1 - 1:a2, b1 - 1:b2] = sub
    return type_map, starts, ends


def find_derivative_regions(flow_region, meshes):
    """Per-direction line type maps (1-based type ids) and decentred-stencil positions.

    Returns ``{'x': (typeMapX[ny,nz], starts, ends), 'y': (typeMapY[nx,nz], ...),
    'z': (typeMapZ[nx,ny], ...)}``.
    """
    fr = np.asarray(flow_region, dtype=bool)
    out = {}
    # x lines: permute(flowRegion,[2 1 3]) -> (ny, nx, nz)
    out["x"] = _regions_one_dir(fr.transpose(1, 0, 2), meshes[0].periodic, fr.shape[0], meshes[0].break_points)
    # y lines: flowRegion itself (nx, ny, nz)
    out["y"] = _regions_one_dir(fr, meshes[1].periodic, fr.shape[1], meshes[1].break_points)
    # z lines: permute(flowRegion,[1 3 2]) -> (nx, nz, ny)
    out["z"] = _regions_one_dir(fr.transpose(0, 2, 1), meshes[2].periodic, fr.shape[2], meshes[2].break_points)
    return out


# ---------------------------------------------------------------------------
# makeMatrices.m
# ---------------------------------------------------------------------------

def make_matrices_each_direction(c_lhs, c_rhs, d_lhs, d_rhs, starts, ends, n, buffer=None):
    """Dense LHS/RHS per line type (``makeMatrices.m:158-263``)."""
    c_lhs = np.atleast_1d(np.asarray(c_lhs, float))
    c_rhs = np.atleast_1d(np.asarray(c_rhs, float))
    d_lhs = np.atleast_2d(np.asarray(d_lhs, float))
    d_rhs = np.atleast_2d(np.asarray(d_rhs, float))
    if n != 1 and n < 2 * len(c_rhs) - 1:
        raise ValueError("Mesh is not large enough for one of the stencils. It has %d nodes but the "
                         "stencil needs at least %d." % (n, 2 * len(c_rhs) - 1))
    ntypes = len(starts)
    if n == 1:
        v = 0.0 if c_rhs[0] == 0 else 1.0
        return [np.ones((1, 1)) for _ in range(ntypes)], [np.full((1, 1), v) for _ in range(ntypes)]
    lhs_base = np.diag(np.full(n, c_lhs[0]))
    rhs_base = np.diag(np.full(n, c_rhs[0]))
    invert = -1.0 if c_rhs[0] == 0 else 1.0
    r = np.arange(n)
    for i in range(1, len(c_lhs)):
        lhs_base[r, (r + i) % n] = c_lhs[i]
        lhs_base[r, (r - i) % n] = c_lhs[i]
    for i in range(1, len(c_rhs)):
        rhs_base[r, (r + i) % n] = c_rhs[i]
        rhs_base[r, (r - i) % n] = invert * c_rhs[i]
    starts = [list(s) for s in starts]
    ends = [list(e) for e in ends]
    if buffer is not None:                             # makeMatrices.m:214-226
        bi, bf = buffer
        if bi.upwind:
            for t in range(ntypes):
                starts[t] = sorted(set(starts[t]) | set(range(1, bi.n + 1)))[::-1]
        if bf.upwind:
            for t in range(ntypes):
                ends[t] = sorted(set(ends[t]) | set(n + 1 - np.arange(1, bf.n + 1)))
    m_l, n_l = d_lhs.shape
    m_r, n_r = d_rhs.shape
    lhs_out, rhs_out = [], []
    for t in range(ntypes):
        L = lhs_base.copy()
        R = rhs_base.copy()
        for s in starts[t]:
            s0 = s - 1
            L[s0:s0 + m_l, :] = 0
            R[s0:s0 + m_r, :] = 0
            L[s0:s0 + m_l, s0:s0 + n_l] = d_lhs
            R[s0:s0 + m_r, s0:s0 + n_r] = d_rhs
        for e in ends[t]:
            e1 = e                                       # exclusive end, 0-based
            L[e1 - m_l:e1, :] = 0
            R[e1 - m_r:e1, :] = 0
            L[e1 - m_l:e1, e1 - n_l:e1] = d_lhs[::-1, ::-1]
            R[e1 - m_r:e1, e1 - n_r:e1] = invert * d_rhs[::-1, ::-1]
        lhs_out.append(L)
        rhs_out.append(R)
    return lhs_out, rhs_out


def add_buffer_to_matrix(base_l, buf_l, base_r, buf_r, n, buffer):
    """Blend the main and buffer schemes with ``sqrt`` of a cosine ramp (``makeMatrices.m:265-311``)."""
    bi, bf = buffer
    ni, nf = bi.n, bf.n
    nti = int(np.round(bi.transition * ni)) if bi.transition is not None else ni
    ntf = int(np.round(bf.transition * nf)) if bf.transition is not None else nf
    ni1, ni2 = ni - nti + 1, ni
    nf1 = n - nf + 1
    nf2 = nf1 + ntf - 1
    eta = np.ones(n)
    if ni > 0:
        eta[0:ni1] = 0
        eta[ni1 - 1:ni2] = 0.5 - 0.5 * np.cos(matlab_linspace(0, np.pi, ni2 - ni1 + 1))
    if nf > 0:
        eta[nf2 - 1:n] = 0
        eta[nf1 - 1:nf2] = 0.5 + 0.5 * np.cos(matlab_linspace(0, np.pi, nf2 - nf1 + 1))
    eta = np.sqrt(eta)[:, None]
    new_l = [eta * bl + (1 - eta) * ul for bl, ul in zip(base_l, buf_l)]
    new_r = [eta * br + (1 - eta) * ur for br, ur in zip(base_r, buf_r)]
    return new_l, new_r


def metric_dxdeta(X, dmesh, method):
    """``dX/deta`` with a compact scheme on the whole line (``makeMatrices.m:313-335``)."""
    X = np.asarray(X, float)
    added = False
    if dmesh.periodic and dmesh.fix_periodic_domain_size and len(X) > 1:
        X = np.concatenate([X, [dmesh.xf]])
        added = True
    n = len(X)
    if n == 1:
        return None
    cl, cr, dl, dr = finite_difference_coefficients(method)
    L, R = make_matrices_each_direction(cl, cr, dl, dr, [[1]], [[n]], n, None)
    d = np.linalg.solve(L[0], R[0] @ X)
    if added:
        d = d[:-1]
    return d


def apply_metric(lhs, X, dmesh, method):
    d = metric_dxdeta(X, dmesh, method)
    if d is None:
        return lhs
    return [L * d[None, :] for L in lhs]


@dataclass
class DirMatrices:
    types: np.ndarray
    LHS: list
    RHS: list
    fLHS: list
    fRHS: list


def make_matrices(meshes, flow_region, nm: NumMethods):
    """Per-direction derivative and filter matrices for every line type (``makeMatrices.m:1-156``)."""
    regions = find_derivative_regions(flow_region, meshes)
    nx, ny, nz = meshes[0].n, meshes[1].n, meshes[2].n
    cl, cr, dl, dr = finite_difference_coefficients(nm.spatial_derivs)
    clb, crb, dlb, drb = finite_difference_coefficients(nm.spatial_derivs_buffer)
    fl, fr_, fdl, fdr = spatial_filter_coefficients(nm.filter_strength, nm.filter_borders)
    if nz > 1 and nz < 2 * len(fr_) - 1:                 # makeMatrices.m:15-25
        flz, frz, fdlz, fdrz = [1.0], [1.0], [[1.0]], [[1.0]]
    else:
        flz, frz, fdlz, fdrz = fl, fr_, fdl, fdr
    sz, ez = regions["z"][1], regions["z"][2]
    if nz == 4 and len(sz) == 1 and len(sz[0]) == 0 and len(ez[0]) == 0:   # makeMatrices.m:28-38
        clz, crz = [1.0], [0.0, np.pi / 4]
        clbz, crbz = clz, crz
    else:
        clz, crz, clbz, crbz = cl, cr, clb, crb
    out = {}
    for d, key, dm, n in ((0, "x", meshes[0], nx), (1, "y", meshes[1], ny), (2, "z", meshes[2], nz)):
        tm, st, en = regions[key]
        c_l, c_r = (clz, crz) if key == "z" else (cl, cr)
        c_lb, c_rb = (clbz, crbz) if key == "z" else (clb, crb)
        f_l, f_r, f_dl, f_dr = (flz, frz, fdlz, fdrz) if key == "z" else (fl, fr_, fdl, fdr)
        L, R = make_matrices_each_direction(c_l, c_r, dl, dr, st, en, n, None)
        buf = (dm.buffer_i, dm.buffer_f)
        Lb, Rb = make_matrices_each_direction(c_lb, c_rb, dlb, drb, st, en, n, buf)
        fL, fR = make_matrices_each_direction(f_l, f_r, f_dl, f_dr, st, en, n, None)
        if nm.change_order[d] and n > 1:
            L, R = add_buffer_to_matrix(L, Lb, R, Rb, n, buf)
        method = nm.metric_method
        if key == "z" and nz == 4:
            method = "SL4"
        L = apply_metric(L, dm.X, dm, method)
        borders_on = nm.filter_borders not in (False, "off") if not isinstance(nm.filter_borders, str) \
            else True
        if borders_on and n > 1:                        # makeMatrices.m:83-133
            if not nm.filter_borders_start[d]:
                for i in range(len(fL)):
                    fL[i][0:5, :] = 0
                    fR[i][0:5, :] = 0
                    fL[i][0:5, 0:5] = np.eye(5)
                    fR[i][0:5, 0:5] = np.eye(5)
            if not nm.filter_borders_end[d]:
                for i in range(len(fL)):
                    fL[i][n - 5:, :] = 0
                    fR[i][n - 5:, :] = 0
                    fL[i][n - 5:, n - 5:] = np.eye(5)
                    fR[i][n - 5:, n - 5:] = np.eye(5)
        out[key] = DirMatrices(types=tm, LHS=L, RHS=R, fLHS=fL, fRHS=fR)
    return out


# ---------------------------------------------------------------------------
# prepareThomas.m
# ---------------------------------------------------------------------------

@dataclass
class LineCoeffs:
    """Pre-factored coefficients of one operator (derivative or filter) along one direction.

    Shapes follow the Fortran arrays (column-major there, ``[..., type]`` last here):
      A (N-1, nT), B (N - periodic, nT), C (N-1, nT), D (N, nT) if periodic,
      R (N, 2*nR-1, nT) with ``R[i, nR-1+l, t]`` multiplying ``F[i+l]``.
    """
    A: np.ndarray
    B: np.ndarray
    C: np.ndarray
    D: np.ndarray
    R: np.ndarray
    nR: int
    periodic: bool

    @property
    def ntypes(self):
        return self.R.shape[2]

    @property
    def n(self):
        return self.R.shape[0]


def _full_diag(M, k):
    """``fullDiag`` of prepareThomas.m:164-172: ``D[i] = M[i, (i+k) mod N]`` per type."""
    N = M.shape[0]
    i = np.arange(N)
    return M[i, (i + k) % N, :]


def _band(rhs_list, N, derivative_precedence):
    T = np.stack([np.asarray(r, float) for r in rhs_list], axis=2)
    diag = _full_diag(T, 0)[:, None, :]
    nd = 1
    while True:
        nxt = _full_diag(T, nd)
        prv = _full_diag(T, -nd)
        if derivative_precedence:
            # prepareThomas.m:115 -- `a || b && c` binds as `a || (b && c)`
            cont = np.any(nxt) or (np.any(prv) and not (2 * nd - 1 > N))
        else:
            cont = (np.any(nxt) or np.any(prv)) and not (2 * nd - 1 > N)
        if not cont:
            break
        nd += 1
        diag = np.concatenate([prv[:, None, :], diag, nxt[:, None, :]], axis=1)
    return diag, nd


def _factor(lhs_list, ntypes, periodic, N):
    A = np.zeros((N - 1, ntypes))
    B = np.zeros((N, ntypes))
    C = np.zeros((N - 1, ntypes))
    D = np.zeros((N, ntypes))
    for t in range(ntypes):
        L = lhs_list[t]
        A[:, t] = np.diag(L, -1)
        B[:, t] = np.diag(L)
        C[:, t] = np.diag(L, 1)
    if periodic:                                       # prepareThomas.m:38-77
        A1 = np.array([lhs_list[t][0, N - 1] for t in range(ntypes)])
        Cn = np.array([lhs_list[t][N - 1, 0] for t in range(ntypes)])
        A, B, C = A[1:], B[1:], C[1:]
        for i in range(N - 2):
            C[i] = C[i] / B[i]
            B[i + 1] = B[i + 1] - A[i] * C[i]
        B = 1.0 / B
        A = np.vstack([A1[None, :], A])
        C = np.vstack([Cn[None, :], C])
        for t in range(ntypes):
            D[:, t] = np.linalg.inv(lhs_list[t])[0, :]
    else:                                              # prepareThomas.m:79-95
        for i in range(N - 1):
            C[i] = C[i] / B[i]
            B[i + 1] = B[i + 1] - A[i] * C[i]
        B = 1.0 / B
    return A, B, C, D


def prepare_thomas(dm: DirMatrices, round_text=True):
    """Return ``(deriv, filt)`` :class:`LineCoeffs` for one direction (``prepareThomas.m:1-162``)."""
    N = dm.LHS[0].shape[0]
    periodic = bool(dm.LHS[0][N - 1, 0] != 0) if N > 1 else False
    ntypes = int(dm.types.max())
    A, B, C, D = _factor(dm.LHS, ntypes, periodic, N)
    Af, Bf, Cf, Df = _factor(dm.fLHS, ntypes, periodic, N)
    R, nR = _band(dm.RHS[:ntypes], N, True)
    Rf, nRf = _band(dm.fRHS[:ntypes], N, False)
    rnd = fortran_round if round_text else (lambda v: np.asarray(v, float))
    der = LineCoeffs(rnd(A), rnd(B), rnd(C), rnd(D) if periodic else None, rnd(R), nR, periodic)
    fil = LineCoeffs(rnd(Af), rnd(Bf), rnd(Cf), rnd(Df) if periodic else None, rnd(Rf), nRf, periodic)
    return der, fil
