"""CPU checks of the isothermal boundary-layer case (boundaryLayerIsothermal.m) on the oracle.

The boundary program must impose what the flow type states: inflow u = 1, v = w = 0, E = E0 and
zero-gradient pressure at x_1; outflow P = P0 and null second derivatives; null second derivatives
at the top; no-slip isothermal wall for x >= 0 and free slip (zero-gradient u, v = 0) for x < 0;
the x break point splits the wall-normal line types at the leading edge.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases


@pytest.fixture(scope="module")
def bl():
    return cases.boundary_layer(nx=160, ny=64, nbuf=12)


def test_mesh_and_operators(bl):
    X, Y = np.asarray(bl.X), np.asarray(bl.Y)
    assert np.all(np.diff(X) > 0) and np.all(np.diff(Y) > 0)
    dy = np.diff(Y)[:51 - 1]
    np.testing.assert_allclose(dy[1:] / dy[:-1], 1.02, rtol=1e-9)       # geometric wall stretching
    assert bl.deriv[0].nR == 5 and bl.deriv[1].nR == 5                    # EX4 in the buffers
    assert bl.deriv[0].ntypes == 2                                        # x break point at the leading edge
    assert bl.sfd_type == 2 and bl.sfd_x.max() == pytest.approx(0.02)
    assert np.all(bl.sfd_x[12:-12] == 0) and np.all(bl.sfd_x[0] > 0)      # chi only in the x buffers
    assert not bl.flow_region[:, 0].any() and bl.flow_region[:, 1:].all()


def test_boundary_conditions(bl, oracle_mod):
    o = oracle_mod.Oracle(bl)
    rng = np.random.default_rng(3)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in bl.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    X = np.asarray(bl.X)
    E0 = bl.E0
    gm1 = bl.gamma - 1
    P = gm1 * st["R"] * st["E"]
    # inflow (Dirichlet rows come before the wall's in the list, the wall overrides j = 1)
    assert np.all(st["U"][0, 1:, 0] == 1.0) and np.all(st["V"][0, 1:, 0] == 0.0)
    np.testing.assert_allclose(st["E"][0, 1:, 0], E0, rtol=1e-15)     # E0 as written with %.20f
    # outflow pressure: R from P0 and the incoming E (Dirichlet phase), before the null second
    # derivative of E moves E (applyBoundCond applies Dirichlet, Neumann, corners, then these)
    # (P0 / (gamma - 1) = E0 for R = P0 / (gamma - 1) / E)
    np.testing.assert_allclose(st["R"][-1, 1:-1, 0], E0 / q["E"][-1, 1:-1, 0], rtol=1e-14)
    # wall: no slip for x >= 0, free slip for x < 0, v = 0 and E = E0 everywhere on it
    up = X < 0
    assert np.all(st["U"][~up, 0, 0] == 0.0)
    c = np.asarray(bl.bi.neumann)
    L = len(c)
    ext = sum(c[k] * st["U"][up, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(st["U"][up, 0, 0], ext, rtol=1e-13, atol=1e-15)
    assert np.all(st["V"][:, 0, 0] == 0.0)
    np.testing.assert_allclose(st["E"][:, 0, 0], E0, rtol=1e-15)
    # zero wall-normal pressure gradient (one-sided extrapolation) on the wall
    extP = sum(c[k] * P[1:-1, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(P[1:-1, 0, 0], extP, rtol=1e-12)
    # top: null second derivative of u (2nd-order one-sided, neumann2 coefficients)
    c2 = np.asarray(bl.bi.neumann2)
    sec = sum(c2[k] * st["U"][:, -2 - k, 0] for k in range(len(c2)))
    np.testing.assert_allclose(st["U"][1:-1, -1, 0], sec[1:-1], rtol=1e-12, atol=1e-14)


def test_steps_stay_bounded(bl, oracle_mod):
    o = oracle_mod.Oracle(bl)
    q0 = o.apply_bound_cond(bl.state0, 0.0)
    o.init_mean(q0)
    st, mc, dv = o.run(q0, 0.0, bl.dt, 20)
    assert not dv and np.all(np.isfinite(st["U"]))
    assert np.abs(st["U"]).max() < 1.5


def test_adiabatic_wall(oracle_mod):
    """boundaryLayerAdiabatic.m: the wall carries zero-gradient E (one-sided extrapolation of
    the neumann order) instead of E = E0; no slip and zero wall-normal pressure gradient stay."""
    bl = cases.boundary_layer(nx=160, ny=64, nbuf=12, adiabatic=True)
    o = oracle_mod.Oracle(bl)
    rng = np.random.default_rng(4)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in bl.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    X = np.asarray(bl.X)
    c = np.asarray(bl.bi.neumann)
    L = len(c)
    plate = X >= 0
    extE = sum(c[k] * st["E"][:, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(st["E"][1:-1, 0, 0], extE[1:-1], rtol=1e-13)
    assert not np.allclose(st["E"][1:-1, 0, 0], bl.E0)          # not the isothermal value
    assert np.all(st["U"][plate, 0, 0] == 0.0) and np.all(st["V"][:, 0, 0] == 0.0)
    P = (bl.gamma - 1) * st["R"] * st["E"]
    extP = sum(c[k] * P[1:-1, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(P[1:-1, 0, 0], extP, rtol=1e-12)
    assert np.all(bl.bi.cAdiabatic == 1)


def test_z_symmetry_planes(oracle_mod):
    """boundaryLayerIsothermalSymmZ.m: on both z planes w = 0 and u, v, e, p have zero z gradient
    (one-sided extrapolation of the neumann order)."""
    bl = cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12, z_symmetry=True)
    assert not bl.deriv[2].periodic
    o = oracle_mod.Oracle(bl)
    rng = np.random.default_rng(5)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in bl.state0.items()}
    st = o.apply_bound_cond(q, 0.0)
    c = np.asarray(bl.bi.neumann)
    L = len(c)
    assert np.all(st["W"][:, :, 0] == 0.0) and np.all(st["W"][:, :, -1] == 0.0)
    for k in "UV":
        lo = sum(c[m] * st[k][:, :, 1 + m] for m in range(L))
        hi = sum(c[m] * st[k][:, :, -2 - m] for m in range(L))
        # rows touched later by the wall and inflow conditions are excluded
        np.testing.assert_allclose(st[k][1:-1, 1:-1, 0], lo[1:-1, 1:-1], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(st[k][1:-1, 1:-1, -1], hi[1:-1, 1:-1], rtol=1e-12, atol=1e-14)


def test_pressure_inlet(oracle_mod):
    """boundaryLayerIsothermalPressureInlet.m: P = P0 on the inflow plane (R = P0/(gamma-1)/E) and a
    zero x gradient of P at the outflow (one-sided extrapolation), the reverse of the default case."""
    bl = cases.boundary_layer(nx=160, ny=64, nbuf=12, pressure_inlet=True)
    o = oracle_mod.Oracle(bl)
    rng = np.random.default_rng(6)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in bl.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    gm1 = bl.gamma - 1
    P0 = gm1 * bl.E0
    P = gm1 * st["R"] * st["E"]
    np.testing.assert_allclose(P[0, 1:-1, 0], P0, rtol=1e-14)
    # the pressure extrapolation sets R = P_ext / E with the E of that moment; the null second
    # derivative of E applied afterwards (applyBoundCond's order) then moves E on the plane
    c = np.asarray(bl.bi.neumann)
    ext = sum(c[k] * P[-2 - k, 1:-1, 0] for k in range(len(c)))
    np.testing.assert_allclose(gm1 * st["R"][-1, 1:-1, 0] * q["E"][-1, 1:-1, 0], ext, rtol=1e-12)


def test_free_slip_wall(oracle_mod):
    """boundaryLayerFreeSlip.m: the whole wall carries zero-gradient u, w and p (one-sided
    extrapolation of the neumann order), v = 0 and E = E0 from the inside boxes; no x break
    point; the inflow, outflow and top as in the isothermal case."""
    bl = cases.boundary_layer(nx=160, ny=64, nbuf=12, free_slip=True)
    assert bl.deriv[0].ntypes == 1
    o = oracle_mod.Oracle(bl)
    rng = np.random.default_rng(7)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in bl.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    c = np.asarray(bl.bi.neumann)
    L = len(c)
    ext = sum(c[k] * st["U"][:, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(st["U"][1:-1, 0, 0], ext[1:-1], rtol=1e-13, atol=1e-15)
    assert np.all(st["U"][1:-1, 0, 0] != 0.0)                    # slip: no zero velocity on the plate
    assert np.all(st["V"][:, 0, 0] == 0.0)
    np.testing.assert_allclose(st["E"][1:-1, 0, 0], bl.E0, rtol=1e-15)
    P = (bl.gamma - 1) * st["R"] * st["E"]
    extP = sum(c[k] * P[1:-1, 1 + k, 0] for k in range(L))
    np.testing.assert_allclose(P[1:-1, 0, 0], extP, rtol=1e-12)
    assert np.all(st["U"][0, 1:, 0] == 1.0)
    q0 = o.apply_bound_cond(bl.state0, 0.0)
    o.init_mean(q0)
    st, mc, dv = o.run(q0, 0.0, bl.dt, 5)
    assert not dv and np.all(np.isfinite(st["U"]))


def test_free_slip_periodic_lists():
    """The periodic switches of boundaryLayerFreeSlip.m: x periodic drops the in- and outflow,
    y periodic drops the top and the plate (all flow, no wall boxes)."""
    from gatt_dns_amd.prep.boundaries import boundary_layer_free_slip
    X = np.linspace(-1.0, 5.0, 24)
    Y = np.linspace(0.0, 3.0, 12)
    Z = np.array([0.0])
    fr, walls, corners, bc, bp, hold = boundary_layer_free_slip(X, Y, Z, 2.0, 0.8, x_periodic=True)
    assert bp is None and not hold
    assert not any(d in ("xi", "xf") and t != "neu" for d, t in zip(bc.dir, bc.type))
    assert sorted(set(v for v, d in zip(bc.var, bc.dir) if d == "yi")) == ["e", "p", "u", "v", "w"]
    fr, walls, corners, bc, bp, hold = boundary_layer_free_slip(X, Y, Z, 2.0, 0.8, y_periodic=True)
    assert fr.all() and not any(d in ("yi", "yf") for d in bc.dir)
    fr, walls, corners, bc, bp, hold = boundary_layer_free_slip(X + 2.0, Y, Z, 2.0, 0.8)
    assert hold and not any(t == "dir" and d == "xi" for d, t in zip(bc.dir, bc.type))


INF = np.inf


def cavity_case():
    """The reference's cavity configuration in small (parameters.m:28-44: domain.yi = -D, a cavity
    [x1 x2] x [-D 0] cut into the plate) with a roughness element on the plate downstream.  Every
    interval ends on a mesh node, as the matched fixed points of the reference's mesh put them, so
    the strict (cavity) and inclusive (element) bounds decide which nodes are walls."""
    kw = dict(nx=96, ny=48, nbuf=10, n_below=8, wall_depth=2.0)
    c0 = cases.boundary_layer(**kw)
    X, Y = np.asarray(c0.X), np.asarray(c0.Y)
    return cases.boundary_layer(cavities=[dict(x=(X[30], X[40]), y=(-2.0, 0.0), z=(-INF, INF))],
                                roughness=[dict(x=(X[50], X[53]), y=(0.0, Y[12]), z=(-INF, INF))], **kw)


def test_cavity_and_roughness(oracle_mod):
    """Cavity and roughness element of boundaryLayerIsothermal.m:338-358: the nodes strictly inside
    the cavity's intervals are opened (and the one-node plate row above them, a wall between flow
    on both sides, goes with them), the nodes on or inside the element's are closed.  After the
    boundary conditions every wall node with flow beside it is no-slip with E = E0, and the
    pressure has zero normal gradient (one-sided extrapolation into the flow) on the cavity's
    side walls and floor and on the element's faces."""
    c = cavity_case()
    X, Y = np.asarray(c.X), np.asarray(c.Y)
    fr = c.flow_region[:, :, 0]
    j0 = int(np.argmin(np.abs(Y)))                      # the plate row, Y = 0
    assert Y[j0] == 0.0 and j0 == 8
    ic = np.arange(31, 40)                               # strictly inside X[30] .. X[40]
    jc = np.arange(1, j0)                                # strictly inside -2 .. 0
    ir = np.arange(50, 54)                               # X[50] .. X[53] inclusive
    jr = np.arange(j0, 13)                               # 0 .. Y[12] inclusive
    want = np.zeros_like(fr)
    want[:, j0 + 1:] = True
    want[np.ix_(ic, jc)] = True
    want[ic, j0] = True                                  # the lip row between the cavity and the flow
    want[np.ix_(ir, jr)] = False
    assert np.array_equal(fr, want)
    o = oracle_mod.Oracle(c)
    rng = np.random.default_rng(4)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in c.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    U, V, E = (st[k][:, :, 0] for k in "UVE")
    P = (c.gamma - 1) * st["R"][:, :, 0] * E
    wall = ~fr
    near = np.zeros_like(wall)                            # wall nodes with a flow node beside them
    near[:-1] |= fr[1:]
    near[1:] |= fr[:-1]
    near[:, :-1] |= fr[:, 1:]
    near[:, 1:] |= fr[:, :-1]
    faces = wall & near & (X[:, None] >= 0)
    assert np.all(U[faces] == 0.0) and np.all(V[faces] == 0.0)
    np.testing.assert_allclose(E[faces], c.E0, rtol=1e-15)
    cn = np.asarray(c.bi.neumann)
    L = len(cn)

    def extrap(i, j, di, dj):
        return sum(cn[k] * P[i + (k + 1) * di, j + (k + 1) * dj] for k in range(L))

    checks = [(ic[0] - 1, j, 1, 0) for j in jc[1:]]               # cavity's upstream wall
    checks += [(ic[-1] + 1, j, -1, 0) for j in jc[1:]]            # downstream wall
    checks += [(i, jc[0] - 1, 0, 1) for i in ic[1:-1]]            # floor
    checks += [(i, jr[-1], 0, 1) for i in ir[1:-1]]               # element's top
    checks += [(ir[0], j, -1, 0) for j in jr[1:-1]]               # element's upstream face
    checks += [(ir[-1], j, 1, 0) for j in jr[1:-1]]               # downstream face
    assert len(checks) > 20
    for i, j, di, dj in checks:
        assert P[i, j] == pytest.approx(extrap(i, j, di, dj), rel=1e-12), (i, j, di, dj)
    st, mc, dv = o.run(o.apply_bound_cond(c.state0, 0.0), 0.0, c.dt, 5)
    assert not dv and np.all(np.isfinite(st["U"]))


def test_mesh_fixed_points():
    """meshAddFixedPoints.m: 0 in x and y, the cavities', elements' and fitted disturbances'
    intervals and the fitted tracked points; sorted, unique, finite and inside the domain."""
    from gatt_dns_amd.prep.mesh import mesh_fixed_points
    fx, fy, fz = mesh_fixed_points(
        (0.0, 300.0, -5.0, 30.0, 0.0, 1.0),
        cavities=[dict(x=(200.0, 210.0), y=(-5.0, 0.0), z=(-INF, INF))],
        roughness=[dict(x=(1.0, 1.2), y=(-0.1, 0.0), z=(0.3, 0.8))],
        disturbances=[dict(x=(25.0, 50.0), y=(0.0, 0.0), z=(-INF, INF), fitPoints=False, fitPointsX=True),
                      dict(x=(400.0, 500.0), y=(1.0, 2.0), z=(0.1, 0.2), fitPoints=True)],
        tracked_points=[[7.0, 3.0, 0.5]], fit_tracked_points=True)
    np.testing.assert_array_equal(fx, [0.0, 1.0, 1.2, 7.0, 25.0, 50.0, 200.0, 210.0])
    np.testing.assert_array_equal(fy, [-5.0, -0.1, 0.0, 1.0, 2.0, 3.0])
    np.testing.assert_array_equal(fz, [0.1, 0.2, 0.3, 0.5, 0.8])
    fx, fy, fz = mesh_fixed_points((0.0, 1.0, 0.0, 1.0, 0.0, 1.0), tracked_points=[[0.5, 0.5, 0.5]])
    np.testing.assert_array_equal(fx, [0.0])
    np.testing.assert_array_equal(fz, [])


def test_lid_driven_obstacle(oracle_mod):
    """A roughness block inside the lid-driven cavity (lidDrivenFlow.m:16-25, before the outer
    walls): its nodes are walls, its faces no-slip with zero normal pressure gradient."""
    c = cases.lid_driven(nx=48, ny=40, roughness=[dict(x=(0.35, 0.6), y=(0.3, 0.55), z=(-INF, INF))])
    X, Y = np.asarray(c.X), np.asarray(c.Y)
    fr = c.flow_region[:, :, 0]
    ib = np.flatnonzero((X >= 0.35) & (X <= 0.6))
    jb = np.flatnonzero((Y >= 0.3) & (Y <= 0.55))
    want = np.ones_like(fr)
    want[[0, -1], :] = False
    want[:, [0, -1]] = False
    want[np.ix_(ib, jb)] = False
    assert np.array_equal(fr, want)
    o = oracle_mod.Oracle(c)
    rng = np.random.default_rng(6)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in c.state0.items()}
    q["W"][:] = 0.0
    st = o.apply_bound_cond(q, 0.0)
    U, V = st["U"][:, :, 0], st["V"][:, :, 0]
    assert np.all(U[np.ix_(ib, jb)] == 0.0) and np.all(V[np.ix_(ib, jb)] == 0.0)
    P = (c.gamma - 1) * st["R"][:, :, 0] * st["E"][:, :, 0]
    cn = np.asarray(c.bi.neumann)
    for j in jb[1:-1]:
        assert P[ib[0], j] == pytest.approx(sum(cn[k] * P[ib[0] - 1 - k, j] for k in range(len(cn))), rel=1e-12)
        assert P[ib[-1], j] == pytest.approx(sum(cn[k] * P[ib[-1] + 1 + k, j] for k in range(len(cn))), rel=1e-12)
    for i in ib[1:-1]:
        assert P[i, jb[0]] == pytest.approx(sum(cn[k] * P[i, jb[0] - 1 - k] for k in range(len(cn))), rel=1e-12)
        assert P[i, jb[-1]] == pytest.approx(sum(cn[k] * P[i, jb[-1] + 1 + k] for k in range(len(cn))), rel=1e-12)
