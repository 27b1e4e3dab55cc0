"""Independent pins of oracle functions that no other test fixes.

* turbulenceGenerator (turbulenceGenerator.F90:20-42, in oracle_nseq): the forcing terms
  du_f = 0.25 eps N / E_T u - gamma mean(u) with gamma = 10 eps inject the energy
  sum(u . du_f) = 0.25 eps N - gamma |sum u|^2 / N, and de_f is the constant
  -gamma (mean(E) - e0); R is not forced.  Checked on a zero-mean and a shifted field.
* the SFD blend (main.F90:381-402, oracle_sfd): per point the pair (q, mean) follows the
  linear map [[1 - a + a b, a (1 - b)], [b, 1 - b]] with a = dt chi, b = dt / Delta
  (b = 0 for a frozen mean); fifty oracle steps equal that 2x2 matrix to the fiftieth
  power (numpy), and points with chi = 0 keep q.
* the filter's blend and direction composition (derivs.F90:399-552, oracle_apply_filters):
  on a periodic box a tensor-product Fourier mode is multiplied by
  prod_d (1 + theta (T_d - 1)), T_d the transfer function of the compact filter at the
  mode's wavenumber from the stencil's own coefficients; theta <= 0 gives prod_d T_d.
* calcSFDregion (calcSFDregion.m:4-62): the ramp is chi cos^2(pi (i - 1) / (2 n)) at buffer
  index i of n, 0 in the physical region, 1 - product of the ramps' complements at a
  corner; an extra region adds X at its centre, X / 2 at half its ellipsoidal radius and
  nothing outside it.
"""
import copy

import numpy as np
import pytest

from gatt_dns_amd.prep import cases
from gatt_dns_amd.prep.boundaries import calc_sfd_region
from gatt_dns_amd.prep.stencils import spatial_filter_coefficients


@pytest.mark.parametrize("shift", [0.0, 0.3])
def test_turbulence_forcing_closed_forms(oracle_mod, shift):
    case = cases.turbulence_box(12)
    eps, e0 = case.forcing[0].par
    gamma = 10 * eps
    off = copy.copy(case)
    off.forcing = []
    rng = np.random.default_rng(4)
    q = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
    q["U"] = q["U"] + shift
    q["E"] = np.asfortranarray(q["E"] * (1 + 0.01 * rng.standard_normal(q["E"].shape)))
    r_on = oracle_mod.Oracle(case).rhs(q)
    r_off = oracle_mod.Oracle(off).rhs(q)
    du = {k: r_on[k] - r_off[k] for k in "UVWRE"}
    n = q["U"].size
    inj = sum(np.sum(q[k] * du[k]) for k in "UVW")
    sums = np.array([np.sum(q[k]) for k in "UVW"])
    want = 0.25 * eps * n - gamma * np.sum(sums ** 2) / n
    assert abs(inj - want) <= 1e-9 * max(1.0, abs(want)), (inj, want)
    np.testing.assert_allclose(du["E"], -gamma * (np.mean(q["E"]) - e0), rtol=1e-9, atol=1e-12)
    assert np.max(np.abs(du["R"])) == 0.0


@pytest.mark.parametrize("delta", [10.0, -1.0])
def test_sfd_blend_is_the_linear_map(oracle_mod, delta):
    case = copy.copy(cases.boundary_layer(nx=96, ny=48, nbuf=10))
    case.sfd_delta = delta
    dt = 0.5
    rng = np.random.default_rng(8)
    shape = (case.nx, case.ny, case.nz)
    q = {k: np.asfortranarray(1.0 + rng.random(shape)) for k in "UVWRE"}
    m = {k: np.asfortranarray(1.0 + rng.random(shape)) for k in "UVWRE"}
    o = oracle_mod.Oracle(case)
    o.init_mean(m)
    steps = 50
    st = q
    for _ in range(steps):
        st = o.sfd(st, dt)
    mean = o.mean()
    chi = np.asarray(case.sfd_x).reshape(shape, order="F")
    a = dt * chi
    b = dt / delta if delta > 0 else 0.0
    # the 2x2 map per point, raised to the 50th power by repeated squaring (numpy)
    A = np.stack([np.stack([1 - a + a * b, a * (1 - b)], -1), np.stack([np.full_like(a, b), np.full_like(a, 1 - b)], -1)],
                 -2)
    P = np.linalg.matrix_power(A, steps)
    for k in "UVRE":
        want_q = P[..., 0, 0] * q[k] + P[..., 0, 1] * m[k]
        want_m = P[..., 1, 0] * q[k] + P[..., 1, 1] * m[k]
        np.testing.assert_allclose(st[k], want_q, rtol=1e-12)
        np.testing.assert_allclose(mean[k], want_m, rtol=1e-12)
        free = chi == 0
        assert free.any()
        np.testing.assert_array_equal(st[k][free], q[k][free])


def transfer(alpha, kh):
    """Transfer function of the interior 10th-order compact filter at wavenumber kh."""
    lhs, rhs, _, _ = spatial_filter_coefficients(alpha, "reducedOrder")
    num = rhs[0] + 2 * sum(rhs[j] * np.cos(j * kh) for j in range(1, len(rhs)))
    return num / (lhs[0] + 2 * lhs[1] * np.cos(kh))


@pytest.mark.parametrize("theta", [0.3, -1.0])
def test_filter_blend_and_composition(oracle_mod, theta):
    nx, ny, nz = 16, 12, 14
    case = cases.periodic_rest(nx, ny, nz)
    modes = {"U": (3, 2, 1), "V": (5, 1, 4), "W": (1, 4, 2), "R": (7, 5, 6), "E": (2, 3, 5)}
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[None, None, :]
    q = {}
    for f, (a, b, c) in modes.items():
        q[f] = np.asfortranarray(np.cos(2 * np.pi * a * i / nx + 0.3) * np.cos(2 * np.pi * b * j / ny + 0.7) *
                                 np.cos(2 * np.pi * c * k / nz + 0.1))
    got = oracle_mod.Oracle(case).apply_filters(q, theta)
    for f, (a, b, c) in modes.items():
        T = [transfer(0.49, 2 * np.pi * m / n) for m, n in ((a, nx), (b, ny), (c, nz))]
        g = np.prod([1 + theta * (t - 1) for t in T]) if theta > 0 else np.prod(T)
        np.testing.assert_allclose(got[f], g * q[f], rtol=0, atol=1e-13)


def test_sfd_region_ramps_and_extra_region():
    nx, ny, nz, n = 40, 30, 1, 8
    chi = 0.02
    f = calc_sfd_region(nx, ny, nz, ((n, n), (0, n), (0, 0)), chi, apply=(2, 2, 2))
    ramp = chi * np.cos(np.pi * np.arange(n) / (2 * n)) ** 2          # index i = 1..n
    np.testing.assert_allclose(f[:n, ny // 2, 0], ramp, rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(f[::-1][:n, ny // 2, 0], ramp, rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(f[nx // 2, ::-1][:n, 0], ramp, rtol=1e-14, atol=1e-18)
    assert f[n:nx - n, :ny - n].max() == 0.0
    # the corner: chi (1 - (1 - rx)(1 - ry)) with rx, ry the ramps over chi
    cx, cy = ramp[2] / chi, ramp[5] / chi
    assert f[2, ny - 6, 0] == pytest.approx(chi * (1 - (1 - cx) * (1 - cy)), rel=1e-14)
    assert f[4, 3, 0] == pytest.approx(ramp[4], rel=1e-14)          # y start not damped (no buffer)
    X = np.linspace(0.0, 4.0, nx)
    Y = np.linspace(0.0, 3.0, ny)
    Z = np.zeros(1)
    er = {"location": (X[20], Y[15], 0.0), "size": (1.0, 0.5, 1.0), "X": 0.1}
    g = calc_sfd_region(nx, ny, nz, ((n, n), (0, n), (0, 0)), chi, apply=(2, 2, 2), X=X, Y=Y, Z=Z,
                        extra_regions=[er])
    d = g - f
    assert d[20, 15, 0] == pytest.approx(0.1, rel=1e-15)
    R = np.sqrt((X[:, None] - X[20]) ** 2 / 1.0 + (Y[None, :] - Y[15]) ** 2 / 0.25)
    assert np.all(d[..., 0][R >= 1] == 0.0)
    half = np.argmin(np.abs(np.abs(X - X[20]) - 0.5))
    r = abs(X[half] - X[20])
    assert d[half, 15, 0] == pytest.approx(0.1 * (0.5 + 0.5 * np.cos(np.pi * r)), rel=1e-14)
    assert abs(r - 0.5) < 0.06 and d[half, 15, 0] == pytest.approx(0.05, abs=0.02)
