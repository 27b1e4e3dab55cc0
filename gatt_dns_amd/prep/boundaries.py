"""Boundary-condition discovery and the per-variable box lists.

Host-side restatement of the reference's boundary preprocessing for the flow
types the time-advance configurations use:

  * ``source/findWallsForBoundaries.m`` -> :func:`find_walls_for_boundaries`
  * ``source/boundaries/lidDrivenFlow.m`` -> :func:`lid_driven_flow`
  * ``source/boundaries/periodicBox.m`` -> :func:`periodic_box`
  * ``source/boundaries/boundaryLayerIsothermal.m`` (and its Adiabatic, SymmZ and
    PressureInlet variants) -> :func:`boundary_layer_isothermal`
  * ``source/boundaries/boundaryLayerFreeSlip.m`` -> :func:`boundary_layer_free_slip`
  * the cavities and roughness elements every flow type cuts into its flow region
    (``flowType.cav`` / ``flowType.rug``) -> :func:`cut_cavities_and_roughness`
  * ``source/initBoundaries.m`` (one rank) -> :func:`init_boundaries`
  * ``source/calcSFDregion.m`` -> :func:`calc_sfd_region`

All index boxes are 1-based and inclusive, ``[i1 i2 j1 j2 k1 k2]``, as in the
generated ``boundaryInfo.F90``.
"""
from dataclasses import dataclass, field

import numpy as np

from .stencils import neumann_coefficients

DIRECTIONS = ("xi", "xf", "yi", "yf", "zi", "zf")     # initBoundaries.m:29


def _runs(mask):
    """1-based (start, end) of the runs of True in a 1-D mask."""
    m = np.asarray(mask, dtype=bool)
    if not m.any():
        return [], []
    d = np.diff(m.astype(np.int8))
    starts = list(np.nonzero(np.concatenate([[m[0]], d == 1]))[0] + 1)
    ends = list(np.nonzero(np.concatenate([d == -1, [m[-1]]]))[0] + 1)
    return starts, ends


def _xyz(el):
    """(x, y, z) pairs of one cavity or roughness element (a dict or an object with x, y, z)."""
    get = (lambda k: el[k]) if isinstance(el, dict) else (lambda k: getattr(el, k))
    return [tuple(float(v) for v in get(k)) for k in "xyz"]


def cut_cavities_and_roughness(fr, X, Y, Z, cavities=(), roughness=()):
    """Open the cavities into the walls of a flow region and close the roughness elements, in
    the order and with the bounds every flow type uses (``boundaryLayerIsothermal.m:338-358``,
    ``lidDrivenFlow.m:5-25``, ``poiseuilleFlow.m:5-25``, ``boundaryLayerFreeSlip.m:346-366``):
    a cavity sets the nodes strictly inside its x, y and z intervals to flow, a roughness
    element sets the nodes inside or on its intervals to wall.  Each element has x, y and z
    pairs (infinite ends allowed).  Modifies and returns ``fr`` (nx x ny x nz)."""
    X, Y, Z = (np.asarray(a, dtype=float) for a in (X, Y, Z))
    for el in cavities:
        x, y, z = _xyz(el)
        fr[np.ix_((X > x[0]) & (X < x[1]), (Y > y[0]) & (Y < y[1]), (Z > z[0]) & (Z < z[1]))] = True
    for el in roughness:
        x, y, z = _xyz(el)
        fr[np.ix_((X >= x[0]) & (X <= x[1]), (Y >= y[0]) & (Y <= y[1]), (Z >= z[0]) & (Z <= z[1]))] = False
    return fr


def _need_xyz(X, Y, Z, cavities, roughness):
    if (cavities or roughness) and (X is None or Y is None or Z is None):
        raise ValueError("cavities and roughness elements need the mesh coordinates X, Y, Z")


def find_walls_for_boundaries(flow_region):
    """Walls, corners and inside-wall boxes of a flow-region mask (``findWallsForBoundaries.m:1-329``).

    Returns ``(flow_region, walls, corners, inside_walls)`` where ``walls`` maps
    front/back/up/down/right/left to lists of boxes and ``corners`` is a dict with
    ``limits``, ``dir`` and ``adiabatic``.
    """
    fr = np.array(flow_region, dtype=bool)
    nx, ny, nz = fr.shape
    # remove infinitely thin walls (lines 7-18)
    if nx > 2:
        fr[1:-1] = fr[1:-1] | (fr[:-2] & fr[2:])
    if ny > 2:
        fr[:, 1:-1] = fr[:, 1:-1] | (fr[:, :-2] & fr[:, 2:])
    if nz > 2:
        fr[:, :, 1:-1] = fr[:, :, 1:-1] | (fr[:, :, :-2] & fr[:, :, 2:])
    fi = fr.astype(np.int8)
    front = np.zeros_like(fr); front[:-1] = np.diff(fi, axis=0) == 1
    back = np.zeros_like(fr); back[1:] = np.diff(fi, axis=0) == -1
    up = np.zeros_like(fr); up[:, :-1] = np.diff(fi, axis=1) == 1
    down = np.zeros_like(fr); down[:, 1:] = np.diff(fi, axis=1) == -1
    right = np.zeros_like(fr); right[:, :, :-1] = np.diff(fi, axis=2) == 1
    left = np.zeros_like(fr); left[:, :, 1:] = np.diff(fi, axis=2) == -1

    def planes(wall, axis):
        """Wall boxes on planes normal to ``axis`` (lines 36-162)."""
        out = []
        if axis == 0:      # loop i, runs along y, merged along z
            for i in range(nx):
                lw = wall[i]
                if not lw.any():
                    continue
                ws = []
                for k in range(nz):
                    if k == 0 or not ws:
                        ws, we = _runs(lw[:, k])
                        kstart = k + 1
                    if k == nz - 1 or np.any(lw[:, k] != lw[:, k + 1]):
                        for a, b in zip(ws, we):
                            out.append([i + 1, i + 1, a, b, kstart, k + 1])
                        ws, we = [], []
        elif axis == 1:    # loop j, runs along x, merged along z
            for j in range(ny):
                lw = wall[:, j, :]
                if not lw.any():
                    continue
                ws = []
                for k in range(nz):
                    if k == 0 or not ws:
                        ws, we = _runs(lw[:, k])
                        kstart = k + 1
                    if k == nz - 1 or np.any(lw[:, k] != lw[:, k + 1]):
                        for a, b in zip(ws, we):
                            out.append([a, b, j + 1, j + 1, kstart, k + 1])
                        ws, we = [], []
        else:              # loop k, runs along y, merged along x
            for k in range(nz):
                lw = wall[:, :, k]
                if not lw.any():
                    continue
                ws = []
                for i in range(nx):
                    if i == 0 or not ws:
                        ws, we = _runs(lw[i, :])
                        istart = i + 1
                    if i == nx - 1 or np.any(lw[i, :] != lw[i + 1, :]):
                        for a, b in zip(ws, we):
                            out.append([istart, i + 1, a, b, k + 1, k + 1])
                        ws, we = [], []
        return out

    walls = {"front": planes(front, 0), "back": planes(back, 0), "up": planes(up, 1),
             "down": planes(down, 1), "right": planes(right, 2), "left": planes(left, 2)}

    # join adjacent walls (lines 164-190)
    for key in ("front", "back", "up", "down"):
        cw = [list(r) for r in walls[key]]
        while True:
            to_merge = None
            for i in range(len(cw) - 1):
                for j in range(i + 1, len(cw)):
                    if cw[i][0:4] == cw[j][0:4] and cw[i][5] == cw[j][4] - 1:
                        to_merge = (i, j)
                if to_merge is not None:
                    break
            if to_merge is None:
                break
            i, j = to_merge
            cw[i][5] = cw[j][5]
            del cw[j]
        walls[key] = cw

    # corners (lines 192-284)
    limits, dirs = [], []
    cm = [front & up, front & down, back & up, back & down]
    cdir = [[1, 1, 0], [1, -1, 0], [-1, 1, 0], [-1, -1, 0]]
    for i in range(nx):
        for j in range(ny):
            for m in range(4):
                s, e = _runs(cm[m][i, j, :])
                for a, b in zip(s, e):
                    limits.append([i + 1, i + 1, j + 1, j + 1, a, b]); dirs.append(cdir[m])
    cm = [front & right, front & left, back & right, back & left]
    cdir = [[1, 0, 1], [1, 0, -1], [-1, 0, 1], [-1, 0, -1]]
    for i in range(nx):
        for k in range(nz):
            for m in range(4):
                s, e = _runs(cm[m][i, :, k])
                for a, b in zip(s, e):
                    limits.append([i + 1, i + 1, a, b, k + 1, k + 1]); dirs.append(cdir[m])
    cm = [up & right, up & left, down & right, down & left]
    cdir = [[0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1]]
    for j in range(ny):
        for k in range(nz):
            for m in range(4):
                s, e = _runs(cm[m][:, j, k])
                for a, b in zip(s, e):
                    limits.append([a, b, j + 1, j + 1, k + 1, k + 1]); dirs.append(cdir[m])
    cm3 = [front & up & right, front & up & left, front & down & right, front & down & left,
           back & up & right, back & up & left, back & down & right, back & down & left]
    cdir3 = [[1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1], [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1]]
    for m in range(8):                      # column-major find over (i,j,k,m)
        idx = np.argwhere(cm3[m].transpose(2, 1, 0))
        for k, j, i in idx:
            limits.append([i + 1, i + 1, j + 1, j + 1, k + 1, k + 1]); dirs.append(cdir3[m])
    corners = {"limits": limits, "dir": dirs, "adiabatic": [0] * len(limits)}

    # regions completely inside walls (lines 286-329)
    is_wall = ~fr
    inside = []
    for k in range(nz):
        for j in range(ny):
            s, e = _runs(is_wall[:, j, k])
            for a, b in zip(s, e):
                inside.append([a, b, j + 1, j + 1, k + 1, k + 1])
    if inside:
        keys, order = [], []
        for r in inside:
            key = (r[0], r[1])
            if key not in keys:
                keys.append(key)
            order.append(keys.index(key))
        perm = sorted(range(len(inside)), key=lambda t: order[t])   # stable
        inside = [inside[t] for t in perm]
    i = 0
    while i < len(inside) - 1:
        a, b = inside[i], inside[i + 1]
        if [a[0], a[1], a[4], a[5]] == [b[0], b[1], b[4], b[5]] and a[3] + 1 == b[2]:
            a[3] = b[3]
            del inside[i + 1]
        else:
            i += 1
    i = 0
    while i < len(inside) - 1:
        a, b = inside[i], inside[i + 1]
        if a[0:4] == b[0:4] and a[5] + 1 == b[4]:
            a[5] = b[5]
            del inside[i + 1]
        else:
            i += 1
    return fr, walls, corners, inside


@dataclass
class BCList:
    """Raw boundary list in the order the flow-type script appends it."""
    var: list = field(default_factory=list)     # 'u','v','w','p','e'
    type: list = field(default_factory=list)    # 'dir','neu','sec'
    dir: list = field(default_factory=list)     # 'xi'..'zf'
    val: list = field(default_factory=list)
    box: list = field(default_factory=list)     # [xi xf yi yf zi zf]

    def add(self, var, typ, d, val, box):
        self.var.append(var); self.type.append(typ); self.dir.append(d)
        self.val.append(float(val)); self.box.append([int(v) for v in box])


def lid_driven_flow(nx, ny, nz, E0, P0, t_wall_relative=None, X=None, Y=None, Z=None, cavities=(),
                    roughness=()):
    """BC list of ``source/boundaries/lidDrivenFlow.m``; cavities and roughness elements (obstacles
    in the cavity) as :func:`cut_cavities_and_roughness`, before the outer walls (``:5-29``)."""
    fr = np.ones((nx, ny, nz), dtype=bool)
    _need_xyz(X, Y, Z, cavities, roughness)
    if cavities or roughness:
        cut_cavities_and_roughness(fr, X, Y, Z, cavities, roughness)
    fr[[0, -1], :, :] = False
    fr[:, [0, -1], :] = False
    fr, walls, corners, inside = find_walls_for_boundaries(fr)
    bc = BCList()
    for key, d in (("front", "xi"), ("back", "xf"), ("up", "yi"), ("down", "yf"), ("right", "zi"), ("left", "zf")):
        for box in walls[key]:
            bc.add("p", "neu", d, 0.0, box)
    e_wall = E0 * t_wall_relative if t_wall_relative is not None else E0
    for box in inside:
        bc.add("p", "dir", "yi", P0, box)
        bc.add("u", "dir", "yi", 0.0, box)
        bc.add("v", "dir", "yi", 0.0, box)
        bc.add("w", "dir", "yi", 0.0, box)
        bc.add("e", "dir", "yi", e_wall, box)
    return fr, walls, corners, bc


def poiseuille_flow(nx, ny, nz, E0, P0, lower_wall_velocity=0.0, upper_wall_velocity=0.0, t_wall_relative=None,
                    X=None, Y=None, Z=None, cavities=(), roughness=()):
    """BC list of ``source/boundaries/poiseuilleFlow.m``: the first and last y rows are walls, with
    moving-wall U values (``:148-168``); cavities and roughness elements as
    :func:`cut_cavities_and_roughness`, before the outer walls (``:5-28``)."""
    fr = np.ones((nx, ny, nz), dtype=bool)
    _need_xyz(X, Y, Z, cavities, roughness)
    if cavities or roughness:
        cut_cavities_and_roughness(fr, X, Y, Z, cavities, roughness)
    fr[:, [0, -1], :] = False
    fr, walls, corners, inside = find_walls_for_boundaries(fr)
    bc = BCList()
    for key, d in (("front", "xi"), ("back", "xf"), ("up", "yi"), ("down", "yf"), ("right", "zi"), ("left", "zf")):
        for box in walls[key]:
            bc.add("p", "neu", d, 0.0, box)
    e_wall = E0 * t_wall_relative if t_wall_relative is not None else E0
    for box in inside:
        bc.add("p", "dir", "yi", P0, box)
        bc.add("u", "dir", "yi", 0.0, box)
        bc.add("v", "dir", "yi", 0.0, box)
        bc.add("w", "dir", "yi", 0.0, box)
        bc.add("e", "dir", "yi", e_wall, box)
    bc.add("u", "dir", "yi", lower_wall_velocity, [1, nx, 1, 1, 1, nz])
    bc.add("u", "dir", "yf", upper_wall_velocity, [1, nx, ny, ny, 1, nz])
    return fr, walls, corners, bc


def boundary_layer_isothermal(X, Y, Z, E0, P0, z_periodic=True, t_wall_relative=None, adiabatic=False,
                              z_symmetry=False, pressure_inlet=False, cavities=(), roughness=()):
    """BC list of ``source/boundaries/boundaryLayerIsothermal.m``, with the cavities and roughness
    elements of ``:338-358`` cut into the plate (:func:`cut_cavities_and_roughness`);
    ``adiabatic=True`` gives ``boundaryLayerAdiabatic.m``: zero-gradient E on every wall
    box beside the pressure condition, adiabatic corners (``:363``), E0 inside the walls,
    and that file's z wall limits (``:387-392``: the left limits carry 'zi', the right 'zf').
    ``z_symmetry=True`` gives the z ends of ``boundaryLayerIsothermalSymmZ.m:205-326``:
    zero-gradient u, v, e, p and w = 0 on both z planes.  ``pressure_inlet=True`` gives the
    pressure conditions of ``boundaryLayerIsothermalPressureInlet.m``: P = P0 at the inflow
    (``:53-64``) and zero-gradient P at the outflow (``:130-140``) in place of the reverse.

    Returns ``(flow_region, walls, corners, bc, x_break_point, hold_inlet)``:
    ``x_break_point`` is the row ``[position, a1, a2, b1, b2]`` of ``mesh.x.breakPoint``
    (``:363-369``) or None, ``hold_inlet`` is True when the domain starts
    downstream of the leading edge and the inflow is held (``:53-67``).
    """
    nx, ny, nz = len(X), len(Y), len(Z)
    bc = BCList()
    full = lambda i1, i2, j1, j2, k1, k2: [i1, i2, j1, j2, k1, k2]
    hold = False
    if X[0] <= 0:                                         # inflow, :4-52
        for var, val in (("u", 1.0), ("v", 0.0), ("w", 0.0), ("e", E0)):
            bc.add(var, "dir", "xi", val, full(1, 1, 1, ny, 1, nz))
        if pressure_inlet:
            bc.add("p", "dir", "xi", P0, full(1, 1, 1, ny, 1, nz))
    else:
        hold = True
    if not pressure_inlet:
        bc.add("p", "neu", "xi", 0.0, full(1, 1, 1, ny, 1, nz))
    for var in "uvwe":                                    # outflow, :82-141
        bc.add(var, "sec", "xf", 0.0, full(nx, nx, 1, ny, 1, nz))
    if pressure_inlet:
        bc.add("p", "neu", "xf", 0.0, full(nx, nx, 1, ny, 1, nz))
    else:
        bc.add("p", "dir", "xf", P0, full(nx, nx, 1, ny, 1, nz))
    for var in "uvwep":                                   # top, :143-202
        bc.add(var, "sec", "yf", 0.0, full(1, nx, ny, ny, 1, nz))
    if nz > 1 and z_symmetry:                             # symmetry planes, SymmZ.m:205-326
        for d, k in (("zi", 1), ("zf", nz)):
            for var in "uvwep":
                bc.add(var, "dir" if var == "w" else "neu", d, 0.0, full(1, nx, 1, ny, k, k))
    elif nz > 1 and not z_periodic:                       # spanwise walls, :204-326
        for d, k in (("zi", 1), ("zf", nz)):
            for var in "uvwep":
                bc.add(var, "neu", d, 0.0, full(1, nx, 1, ny, k, k))
    fr = np.ones((nx, ny, nz), dtype=bool)               # :328-332
    wall_j = int(np.argmin(np.abs(np.asarray(Y)))) + 1
    fr[:, :wall_j, :] = False
    cut_cavities_and_roughness(fr, X, Y, Z, cavities, roughness)
    fr, walls, corners, inside = find_walls_for_boundaries(fr)
    bp = None
    if X[0] < 0:                                          # :363-369
        i0 = int(np.flatnonzero(np.asarray(X) >= 0)[0]) + 1
        i_last = int(np.flatnonzero(np.asarray(X) < 0)[-1]) + 1
        j0 = int(np.flatnonzero(np.asarray(Y) >= 0)[0]) + 1
        for w in walls["up"]:
            w[0] = max(w[0], i0)
        walls["up"].append([1, i_last, j0, j0, 1, nz])
        bp = [i_last, j0, j0, 1, nz]
    zwalls = (("left", "zi"), ("right", "zf")) if adiabatic else (("right", "zi"), ("left", "zf"))
    for key, d in (("front", "xi"), ("back", "xf"), ("up", "yi"), ("down", "yf")) + zwalls:
        for box in walls[key]:                            # :371-402
            bc.add("p", "neu", d, 0.0, box)
            if adiabatic:                                 # boundaryLayerAdiabatic.m:407-418
                bc.add("e", "neu", d, 0.0, box)
    if adiabatic:
        corners["adiabatic"] = [1] * len(corners["limits"])
    if X[0] < 0:                                          # free slip upstream of the plate, :404-432
        bc.add("u", "neu", "yi", 0.0, walls["up"][-1])
        bc.add("w", "neu", "yi", 0.0, walls["up"][-1])
    e_wall = E0 * t_wall_relative if t_wall_relative is not None and not adiabatic else E0
    for box in inside:                                    # :434-492
        bc.add("p", "dir", "yi", P0, box)
        bc.add("u", "dir", "yi", 0.0, box)
        bc.add("v", "dir", "yi", 0.0, box)
        bc.add("w", "dir", "yi", 0.0, box)
        bc.add("e", "dir", "yi", e_wall, box)
    return fr, walls, corners, bc, bp, hold


def boundary_layer_free_slip(X, Y, Z, E0, P0, x_periodic=False, y_periodic=False, z_periodic=True,
                             t_wall_relative=None, cavities=(), roughness=()):
    """BC list of ``source/boundaries/boundaryLayerFreeSlip.m``, with its cavities and roughness
    elements (``:346-366``, :func:`cut_cavities_and_roughness`).

    The in- and outflow of the isothermal boundary layer when x is not periodic (``:4-144``,
    held inflow when the domain starts past x = 0), the zero-gradient top when y is not
    periodic (``:146-209``), zero-gradient z ends for a non-periodic 3-D z (``:211-334``) and
    the plate below y = 0 when y is not periodic (``:336-344``).  Its wall faces carry
    zero-gradient u, w and p (``:371-433``: all u boxes, then w, then p, for each face
    direction) and the points inside the walls u = v = w = 0, p = P0, e = eWall
    (``:435-498``); with v = 0 from the inside boxes this is a slip wall.  There is no
    x break point and no separate slip region ahead of the plate.

    Returns ``(flow_region, walls, corners, bc, None, hold_inlet)`` like
    :func:`boundary_layer_isothermal`.
    """
    nx, ny, nz = len(X), len(Y), len(Z)
    bc = BCList()
    full = lambda i1, i2, j1, j2, k1, k2: [i1, i2, j1, j2, k1, k2]
    hold = False
    if not x_periodic:
        if X[0] <= 0:
            for var, val in (("u", 1.0), ("v", 0.0), ("w", 0.0), ("e", E0)):
                bc.add(var, "dir", "xi", val, full(1, 1, 1, ny, 1, nz))
        else:
            hold = True
        bc.add("p", "neu", "xi", 0.0, full(1, 1, 1, ny, 1, nz))
        for var in "uvwe":
            bc.add(var, "sec", "xf", 0.0, full(nx, nx, 1, ny, 1, nz))
        bc.add("p", "dir", "xf", P0, full(nx, nx, 1, ny, 1, nz))
    if not y_periodic:
        for var in "uvwep":
            bc.add(var, "sec", "yf", 0.0, full(1, nx, ny, ny, 1, nz))
    if nz > 1 and not z_periodic:
        for d, k in (("zi", 1), ("zf", nz)):
            for var in "uvwep":
                bc.add(var, "neu", d, 0.0, full(1, nx, 1, ny, k, k))
    fr = np.ones((nx, ny, nz), dtype=bool)
    if not y_periodic:
        wall_j = int(np.argmin(np.abs(np.asarray(Y)))) + 1
        fr[:, :wall_j, :] = False
    cut_cavities_and_roughness(fr, X, Y, Z, cavities, roughness)
    fr, walls, corners, inside = find_walls_for_boundaries(fr)
    for key, d in (("front", "xi"), ("back", "xf"), ("up", "yi"), ("down", "yf"), ("right", "zi"), ("left", "zf")):
        for var in "uwp":
            for box in walls[key]:
                bc.add(var, "neu", d, 0.0, box)
    e_wall = E0 * t_wall_relative if t_wall_relative is not None else E0
    for box in inside:
        bc.add("p", "dir", "yi", P0, box)
        bc.add("u", "dir", "yi", 0.0, box)
        bc.add("v", "dir", "yi", 0.0, box)
        bc.add("w", "dir", "yi", 0.0, box)
        bc.add("e", "dir", "yi", e_wall, box)
    return fr, walls, corners, bc, None, hold


def periodic_box(nx, ny, nz):
    """``source/boundaries/periodicBox.m``: all flow, no walls."""
    fr, walls, corners, inside = find_walls_for_boundaries(np.ones((nx, ny, nz), dtype=bool))
    return fr, walls, corners, BCList()


@dataclass
class BoundaryInfo:
    """Single-rank ``boundaryInfo`` in the layout of ``applyBoundCond`` (nseq.F90:327-350).

    ``d[var]`` = (boxes (n,6) int, values (n,)) for Dirichlet; ``n[var]`` / ``s[var]`` =
    (boxes, directions 1..6) for Neumann and null-second-derivative conditions,
    with ``var`` in 'UVWPE'.  Corners: ``cL`` (cN,6), ``cD`` (cN,3), ``cAdiabatic`` (cN,).
    """
    d: dict
    n: dict
    s: dict
    cL: np.ndarray
    cD: np.ndarray
    cAdiabatic: np.ndarray
    neumann: np.ndarray
    neumann2: np.ndarray


def init_boundaries(bc: BCList, corners, neumann_order, neumann2_order):
    """Group the BC list per variable and type (``initBoundaries.m:18-118``, one rank)."""
    from .matrices import fortran_round
    d = {v: ([], []) for v in "UVWPE"}
    n = {v: ([], []) for v in "UVWPE"}
    s = {v: ([], []) for v in "UVWPE"}
    for var, typ, dr, val, box in zip(bc.var, bc.type, bc.dir, bc.val, bc.box):
        V = var.upper()
        if typ == "dir":
            d[V][0].append(box); d[V][1].append(val)
        elif typ == "neu":
            n[V][0].append(box); n[V][1].append(DIRECTIONS.index(dr) + 1)
        elif typ == "sec":
            s[V][0].append(box); s[V][1].append(DIRECTIONS.index(dr) + 1)
    conv = lambda lst, dt: (np.array(lst[0], dtype=np.int32).reshape(-1, 6), np.array(lst[1], dtype=dt))
    nc1, nc2 = neumann_coefficients(neumann_order, neumann2_order)
    return BoundaryInfo(
        d={v: (np.array(d[v][0], dtype=np.int32).reshape(-1, 6), fortran_round(np.array(d[v][1], float)))
           for v in "UVWPE"},
        n={v: conv(n[v], np.int32) for v in "UVWPE"},
        s={v: conv(s[v], np.int32) for v in "UVWPE"},
        cL=np.array(corners["limits"], dtype=np.int32).reshape(-1, 6),
        cD=np.array(corners["dir"], dtype=np.int32).reshape(-1, 3),
        cAdiabatic=np.array(corners["adiabatic"], dtype=np.int32).reshape(-1),
        neumann=fortran_round(nc1), neumann2=fortran_round(nc2))


def disturbance_index(X, Y, Z, x, y, z, extra=(0, 0, 0, 0, 0, 0)):
    """Index box of a disturbance region (``getBoundaryConditions.m:106-147``)."""
    def clamp(lo, hi, arr):
        lo = arr[0] if np.isinf(lo) and lo < 0 else (arr[-1] if np.isinf(lo) else lo)
        hi = arr[0] if np.isinf(hi) and hi < 0 else (arr[-1] if np.isinf(hi) else hi)
        return lo, hi
    out = []
    for (lo, hi), arr in (((x[0], x[1]), X), ((y[0], y[1]), Y), ((z[0], z[1]), Z)):
        lo, hi = clamp(lo, hi, np.asarray(arr))
        out.append(int(np.nonzero(np.asarray(arr) >= lo)[0][0]) + 1)
        out.append(int(np.nonzero(np.asarray(arr) <= hi)[0][-1]) + 1)
    ex = list(extra)
    return [out[0] - ex[0], out[1] + ex[1], out[2] - ex[2], out[3] + ex[3], out[4] - ex[4], out[5] + ex[5]]


def calc_sfd_region(nx, ny, nz, buffers, chi, apply=(2, 2, 2), X=None, Y=None, Z=None, extra_regions=()):
    """Buffer-zone damping field ``chi(x,y,z)`` (``calcSFDregion.m:4-62``).

    ``buffers`` = ((nxi, nxf), (nyi, nyf), (nzi, nzf)); ``apply`` = applyX/Y/Z with
    2 = both ends, -1 = start only, 1 = end only, 0 = none.  ``extra_regions``
    (``numMethods.SFD.extraRegion``, ``:52-62``): dicts with ``location`` (x, y, z),
    ``size`` (sx, sy, sz) and ``X``; each adds ``X (0.5 + 0.5 cos(pi R))`` with R the
    ellipsoidal distance from the location, capped at 1, on the mesh X, Y, Z.
    """
    base = _sfd_buffers(nx, ny, nz, buffers, chi, apply)
    if not extra_regions:
        return base
    Xc = np.asarray(X, float)[:, None, None]
    Yc = np.asarray(Y, float)[None, :, None]
    Zc = np.asarray(Z, float)[None, None, :]
    for er in extra_regions:
        loc, size = er["location"], er["size"]
        R = np.sqrt((Xc - loc[0]) ** 2 / size[0] ** 2 + (Yc - loc[1]) ** 2 / size[1] ** 2 +
                    (Zc - loc[2]) ** 2 / size[2] ** 2)
        R = np.minimum(R, 1.0)
        base = base + er["X"] * (0.5 + 0.5 * np.cos(np.pi * R))
    return base


def _sfd_buffers(nx, ny, nz, buffers, chi, apply):
    """The cosine ramps of ``calcSFDregion.m:4-50``: 0 in the physical region, chi at the
    outer edge of each buffer, the factors of overlapping ramps multiplied."""
    f = np.ones((nx, ny, nz))
    (xi, xf), (yi, yf), (zi, zf) = buffers
    ax, ay, az = apply
    if ax in (2, -1):
        for i in range(1, xi + 1):
            f[i - 1] *= 0.5 - 0.5 * np.cos(np.pi * (i - 1) / xi)
    if ax in (2, 1):
        for i in range(1, xf + 1):
            f[nx - i] *= 0.5 - 0.5 * np.cos(np.pi * (i - 1) / xf)
    if ay in (2, -1):
        for j in range(1, yi + 1):
            f[:, j - 1] *= 0.5 - 0.5 * np.cos(np.pi * (j - 1) / yi)
    if ay in (2, 1):
        for j in range(1, yf + 1):
            f[:, ny - j] *= 0.5 - 0.5 * np.cos(np.pi * (j - 1) / yf)
    if az in (2, -1):
        for k in range(1, zi + 1):
            f[:, :, k - 1] *= 0.5 - 0.5 * np.cos(np.pi * (k - 1) / zi)
    if az in (2, 1):
        for k in range(1, zf + 1):
            f[:, :, nz - k] *= 0.5 - 0.5 * np.cos(np.pi * (k - 1) / zf)
    return chi * (1 - f)
