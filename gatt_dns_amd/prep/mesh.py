"""Mesh generation of the preprocessing (``source/generateMesh.m``).

One direction at a time: the physical nodes (``uniform``, ``power``, ``tanh`` refined at
the start ``'i'``, the end ``'f'`` or both ``'b'``, ``attractors``), the matching of
fixed points (``matchFixed`` 1 = spline, 2 = pchip correction), the buffer zones
(``exponential`` with or without a ramp, ``sigmoid``), the extra refinement nodes and
the temporary end node of periodic directions with a fixed domain size.  The ``file``
type reads the nodes from a MAT file (the variable named after the direction) or a text
file of numbers.  MATLAB's
``interp1(..., 'spline')`` is the not-a-knot cubic spline and ``'pchip'`` the
shape-preserving piecewise cubic; scipy's ``CubicSpline`` and ``PchipInterpolator``
implement the same interpolants.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy.interpolate import CubicSpline, PchipInterpolator


@dataclass
class BufferSpec:
    """``mesh.buffer.i`` / ``mesh.buffer.f``: n nodes (or a length l to reach), type and stretching."""
    n: int = 0
    l: float = None
    type: str = "sigmoid"
    stretching: float = 0.0
    ramp: float = None


@dataclass
class MeshSpec:
    """The fields of ``mesh.x`` / ``mesh.y`` / ``mesh.z`` that generateMesh reads."""
    n: int = None
    d0: float = None
    type: str = "uniform"
    local: str = "b"
    par: float = 1.0
    power: float = 1.0
    attractor_points: list = field(default_factory=list)
    attractor_strength: list = field(default_factory=list)
    attractor_size: list = field(default_factory=list)
    attractor_regions: list = field(default_factory=list)   # rows [x1 x2 x3 x4 strength]
    match_fixed: int = 0
    fix_points: list = field(default_factory=list)
    periodic: bool = False
    fix_periodic_domain_size: bool = False
    extra_refinement: int = 0
    buffer_i: BufferSpec = field(default_factory=BufferSpec)
    buffer_f: BufferSpec = field(default_factory=BufferSpec)
    file: str = None                 # mesh.file of the 'file' type
    file_calc_buffer: bool = False   # mesh.fileCalcBuffer: the file holds the physical nodes only


def load_mesh_file(path, direction):
    """The node coordinates of a 'file' mesh (generateMesh.m:162-169): a ``.mat`` file's
    variable named ``direction`` ('x', 'y' or 'z'), otherwise a text file of numbers as
    MATLAB's ``load`` reads it; a column becomes a row."""
    if str(path).endswith(".mat"):
        from scipy.io import loadmat
        d = loadmat(path)
        if direction not in d:
            raise ValueError("%s has no variable %r" % (path, direction))
        X = np.asarray(d[direction], dtype=float)
    else:
        X = np.atleast_2d(np.loadtxt(path, dtype=float))
    if X.ndim == 2 and X.shape[1] == 1:
        X = X.T
    return np.asarray(X, dtype=float).ravel()


def calc_buffer_zone(b):
    """``calcBufferZone`` (generateMesh.m:251-288): cumulative buffer offsets in units of the
    spacing next to the buffer."""
    n = int(b.n)
    if b.type == "exponential":
        if b.ramp is None:
            XB = np.zeros(n)
            XB[0] = 1 + b.stretching
            for i in range(2, n + 1):
                XB[i - 1] = XB[i - 2] + (1 + b.stretching) ** i
            return XB
        delta = np.ones(n)
        for i in range(2, n + 1):
            stretching = 1 + min(1.0, i / b.ramp) * b.stretching
            delta[i - 1] = delta[i - 2] * stretching
        return np.cumsum(delta)
    if b.type == "sigmoid":
        i = np.arange(1, n + 1)
        xb = -10 + 12 * (i - 1) / (n - 1)
        delta = 1 / (1 + np.exp(-1.0 * xb))
        delta = delta * b.stretching + 1
        return np.cumsum(delta)
    raise ValueError("unknown buffer type %r" % b.type)


def _physical(xi, xf, m):
    n = m.n
    if m.type == "uniform":
        return np.linspace(xi, xf, n)
    if m.type == "power":
        return xi + (xf - xi) * np.linspace(0, 1, n) ** m.power
    if m.type == "tanh":
        if m.local == "f":
            eta = np.tanh(np.linspace(0, m.par, n))
            eta = eta / eta[-1]
        elif m.local == "i":
            eta = np.tanh(np.linspace(0, m.par, n))
            eta = eta / eta[-1]
            eta = 1 - eta[::-1]
        else:
            eta = np.tanh(np.linspace(-m.par, m.par, n))
            eta = eta / eta[-1]
            eta = (eta + 1) / 2
        return (xf - xi) * eta + xi
    if m.type == "attractors":
        xb = np.linspace(xi, xf, 100 * n)
        eta = np.ones(100 * n)
        for p, s, w in zip(m.attractor_points, m.attractor_strength, m.attractor_size):
            eta = eta + s * np.exp(-((xb - p) / w) ** 2)
        for row in m.attractor_regions:
            pos = list(row[:4])
            if np.isinf(pos[1]):
                pos[0:2] = [xi - 2, xi - 1]
            if np.isinf(pos[2]):
                pos[2:4] = [xf + 1, xf + 2]
            pos = [pos[0] - 1] + pos + [pos[3] + 1]
            eta = eta + row[4] * PchipInterpolator(pos, [0, 0, 1, 1, 0, 0])(xb) ** 2
        eta = np.cumsum(eta)
        eta = eta - eta[0]
        eta = (n - 1) * eta / eta[-1] + 1
        eta[-1] = n
        return CubicSpline(eta, xb)(np.arange(1, n + 1))
    raise ValueError("mesh type %r is not restated" % m.type)


def generate_mesh(xi, xf, m, direction="x"):
    """``generateMesh(xi, xf, mesh, direction)`` (generateMesh.m:1-249) for one direction;
    returns the node coordinates X (buffers included).

    A 'file' mesh (``:162-180``) without ``file_calc_buffer`` is returned as read, after
    checking that it holds n plus the buffer nodes; with it, the file holds the n physical
    nodes and the fixed-point matching, buffers and extra refinement follow as for the other
    types.  A 'file' mesh takes no temporary periodic end node (``:84``)."""
    import copy
    m = copy.deepcopy(m)
    if m.n is None:                                               # :6-19
        t = copy.deepcopy(m)
        t.n = int(np.ceil((xf - xi) / m.d0))
        t.match_fixed = 0
        t.extra_refinement = 0
        t.fix_periodic_domain_size = False
        t.buffer_i, t.buffer_f = BufferSpec(), BufferSpec()
        d_base = np.max(np.diff(generate_mesh(xi, xf, t, direction)))
        m.n = int(np.ceil(t.n * d_base / m.d0))
    for side in ("buffer_i", "buffer_f"):                          # :22-69
        b = getattr(m, side)
        if b.n is None:
            t = copy.deepcopy(m)
            t.match_fixed = 0
            t.extra_refinement = 0
            t.fix_periodic_domain_size = False
            t.buffer_i, t.buffer_f = BufferSpec(), BufferSpec()
            Xt = generate_mesh(xi, xf, t, direction)
            d_base = Xt[1] - Xt[0] if side == "buffer_i" else Xt[-1] - Xt[-2]
            b.n = int(np.floor(b.l / d_base))
            value = calc_buffer_zone(b)[-1] * d_base
            while value > b.l and b.n > 2:
                b.n = b.n // 2
                value = calc_buffer_zone(b)[-1] * d_base
            while value < b.l:
                b.n += 1
                value = calc_buffer_zone(b)[-1] * d_base
    if m.n == 1:
        m.type = "uniform"
    if m.periodic and (m.buffer_i.n > 0 or m.buffer_f.n > 0):    # :77-82
        m.buffer_i.n = m.buffer_f.n = 0
    added = m.periodic and m.fix_periodic_domain_size and m.n > 1 and m.type != "file"
    if added:
        m.n += 1
    ni, nf = m.buffer_i.n, m.buffer_f.n
    nx = m.n + ni + nf
    if m.type == "file":
        Xf = load_mesh_file(m.file, direction)
        if not m.file_calc_buffer:
            if len(Xf) != nx:
                raise ValueError("%s mesh from file %s contains %d nodes instead of %d as specified in parameters"
                                 % (direction, m.file, len(Xf), nx))
            return Xf
        if len(Xf) != m.n:
            raise ValueError("%s mesh from file %s contains %d nodes instead of %d as specified in parameters"
                             % (direction, m.file, len(Xf), m.n))
        Xp = Xf
    else:
        Xp = _physical(xi, xf, m)
    if m.match_fixed and m.n > 1:                                  # :182-204
        fix = np.unique(np.concatenate([Xp[[0, 1]], np.asarray(m.fix_points, float), Xp[[-2, -1]]]))
        closest = np.argmin(np.abs(Xp[:, None] - fix[None, :]), axis=0) + 1     # 1-based
        it = 0
        while np.any(np.diff(closest) == 0) and it <= m.n:
            dup = np.flatnonzero(np.diff(closest) == 0)
            closest[dup] -= 1
            it += 1
        shift = Xp[closest - 1] - fix
        nodes = np.arange(1, m.n + 1)
        corr = PchipInterpolator(closest, shift)(nodes) if m.match_fixed == 2 else CubicSpline(closest, shift)(nodes)
        Xp = Xp - corr
        Xp[closest - 1] = fix
    X = np.zeros(nx)
    X[ni:ni + m.n] = Xp
    if ni > 0:                                                     # :206-220
        XB = calc_buffer_zone(m.buffer_i) * (X[ni + 1] - X[ni])
        X[:ni] = X[ni] - XB[::-1]
    if nf > 0:
        e = ni + m.n - 1
        XB = calc_buffer_zone(m.buffer_f) * (X[e] - X[e - 1])
        X[e + 1:] = X[e] + XB
    if m.extra_refinement > 0:                                     # :222-236
        er = m.extra_refinement
        n2 = (1 + er) * m.n - er + (1 + er) * (ni + nf)
        Xn = PchipInterpolator(np.linspace(0, 1, len(X)), X)(np.linspace(0, 1, n2))
        Xn[::(1 + er)] = X
        X = Xn
    if added:                                                      # :238-243
        X = X[:-1]
    return X


def mesh_fixed_points(domain, cavities=(), roughness=(), disturbances=(), tracked_points=None,
                      fit_tracked_points=False):
    """The fixed points of each direction (``source/meshAddFixedPoints.m:1-52``), which
    :func:`generate_mesh` matches when ``match_fixed`` is set.

    x and y start from 0, z from nothing; every cavity's and roughness element's x, y and z
    intervals are added, then those of the disturbances with ``fitPoints`` (all three
    directions) or ``fitPointsX/Y/Z`` (that direction), then the tracked points when
    ``fit_tracked_points``.  Each list is sorted without repeats, and infinite points and
    points outside ``domain`` = (xi, xf, yi, yf, zi, zf) are dropped.  Elements and
    disturbances are dicts or objects with x, y, z (and the fit flags); returns (fx, fy, fz).
    """
    def get(el, k, default=None):
        if isinstance(el, dict):
            return el.get(k, default)
        return getattr(el, k, default)

    pts = [[0.0], [0.0], []]
    for el in list(cavities) + list(roughness):
        for d, k in enumerate("xyz"):
            pts[d] += [float(v) for v in get(el, k)]
    for el in disturbances:
        if get(el, "fitPoints", False):
            for d, k in enumerate("xyz"):
                pts[d] += [float(v) for v in get(el, k)]
        else:
            for d, k in enumerate("xyz"):
                if get(el, "fitPoints" + k.upper(), False):
                    pts[d] += [float(v) for v in get(el, k)]
    if tracked_points is not None and fit_tracked_points:
        for row in np.atleast_2d(np.asarray(tracked_points, dtype=float)):
            for d in range(3):
                pts[d].append(float(row[d]))
    out = []
    for d in range(3):
        u = np.unique(np.asarray(pts[d], dtype=float))
        lo, hi = domain[2 * d], domain[2 * d + 1]
        out.append(u[~(np.isinf(u) | (u < lo) | (u > hi))])
    return tuple(out)
