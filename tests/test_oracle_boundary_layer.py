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
