"""Compressible boundary-layer initial flow (``source/calcCompressibleBL.m:1-279``).

The similarity solution of the compressible zero-pressure-gradient boundary layer with
a Sutherland viscosity (Chapman-Rubesin factor C = sqrt(T)(1 + S)/(T + S), S = 110 K / T0):

    (C f'')' + f f'' = 0,     (C T' / Pr)' + f T' + (gamma - 1) Ma^2 C f''^2 = 0,

with f(0) = f'(0) = 0, f'(inf) = 1, T(inf) = 1 and T(0) = Twall (isothermal) or
T'(0) = 0 (adiabatic).  As the reference does, it is written as five first-order
equations (f, f', f'', T, T'), integrated with classical RK4 on eta in [0, 10] (500
steps) and shot with a finite-difference Newton iteration on the two unknown wall values
(``solve_compressibleBL``, ``:96-225``; the RK stages of ``:246-279``).  The wall-normal
coordinate is the Howarth transform y = int_0^eta T (trapezoid rule, ``:217-220``).

The flow (``:1-93``) is the profile at the reference station x_R = Re / 1.7208^2 (where
the incompressible displacement thickness is 1) stretched to each x > 0 with
sqrt(x / x_R) and interpolated with MATLAB's ``makima`` (scipy's modified Akima
interpolator), constant beyond the profile's end; x <= 0 keeps the free stream, the node
at x = 0, y = 0 has U = 0, rows below y = 0 copy the row at y = 0.

Bitwise agreement with MATLAB's ``interp1`` is parity unpinned (no MATLAB here); the
solution is pinned by what the equations fix (tests/test_compressible_bl.py).
"""
import numpy as np
from scipy.interpolate import Akima1DInterpolator

SUTHERLAND_K = 110.0
BLASIUS_SCALE = 1.7208


def _rhs3(y1, y3, y4, y5, c):
    return -y3 * ((y5 / (2 * y4)) - (y5 / (y4 + c))) - y1 * y3 * ((y4 + c) / (np.sqrt(y4) * (1 + c)))


def _rhs5(y1, y3, y4, y5, c, ma, pr, gamma):
    return (-y5 ** 2 * ((0.5 / y4) - (1 / (y4 + c))) - pr * y1 * y5 / np.sqrt(y4) * (y4 + c) / (1 + c)
            - (gamma - 1) * pr * ma ** 2 * y3 ** 2)


def _integrate(n, h, y, c, ma, pr, gamma):
    """RK4 over the n steps of h (calcCompressibleBL.m:246-279); y[k][0] holds the wall values."""
    y1, y2, y3, y4, y5 = y
    for i in range(n):
        k11 = y2[i]
        k21 = y3[i]
        k31 = _rhs3(y1[i], y3[i], y4[i], y5[i], c)
        k41 = y5[i]
        k51 = _rhs5(y1[i], y3[i], y4[i], y5[i], c, ma, pr, gamma)
        k12 = y2[i] + 0.5 * h * k21
        k22 = y3[i] + 0.5 * h * k31
        k32 = _rhs3(y1[i] + 0.5 * h * k11, y3[i] + 0.5 * h * k31, y4[i] + 0.5 * h * k41, y5[i] + 0.5 * h * k51, c)
        k42 = y5[i] + 0.5 * h * k51
        k52 = _rhs5(y1[i] + 0.5 * h * k11, y3[i] + 0.5 * h * k31, y4[i] + 0.5 * h * k41, y5[i] + 0.5 * h * k51,
                    c, ma, pr, gamma)
        k13 = y2[i] + 0.5 * h * k22
        k23 = y3[i] + 0.5 * h * k32
        k33 = _rhs3(y1[i] + 0.5 * h * k12, y3[i] + 0.5 * h * k32, y4[i] + 0.5 * h * k42, y5[i] + 0.5 * h * k52, c)
        k43 = y5[i] + 0.5 * h * k52
        k53 = _rhs5(y1[i] + 0.5 * h * k12, y3[i] + 0.5 * h * k32, y4[i] + 0.5 * h * k42, y5[i] + 0.5 * h * k52,
                    c, ma, pr, gamma)
        k14 = y2[i] + h * k23
        k24 = y3[i] + h * k33
        k34 = _rhs3(y1[i] + h * k13, y3[i] + h * k33, y4[i] + h * k43, y5[i] + h * k53, c)
        k44 = y5[i] + h * k53
        k54 = _rhs5(y1[i] + h * k13, y3[i] + h * k33, y4[i] + h * k43, y5[i] + h * k53, c, ma, pr, gamma)
        y5[i + 1] = y5[i] + (1 / 6) * (k51 + 2 * k52 + 2 * k53 + k54) * h
        y4[i + 1] = y4[i] + (1 / 6) * (k41 + 2 * k42 + 2 * k43 + k44) * h
        y3[i + 1] = y3[i] + (1 / 6) * (k31 + 2 * k32 + 2 * k33 + k34) * h
        y2[i + 1] = y2[i] + (1 / 6) * (k21 + 2 * k22 + 2 * k23 + k24) * h
        y1[i + 1] = y1[i] + (1 / 6) * (k11 + 2 * k12 + 2 * k13 + k14) * h


def solve_compressible_bl(ma, pr, t_inf, t_wall, gamma, adiabatic, lim=10.0, n=500, delta=1e-10, tol=1e-9,
                          max_iter=100000):
    """The similarity solution (solve_compressibleBL, calcCompressibleBL.m:96-225).

    Returns a dict with eta (the Howarth-transformed wall-normal coordinate), f, f_p,
    f_pp, rbar (T / T_inf) and rbar_p, and ``iterations``.  As in the reference, the
    returned profile is the last of the three integrations of the final Newton step.
    """
    c = SUTHERLAND_K / t_inf
    h = lim / n
    eta = np.arange(n + 1) * h
    y = [np.zeros(n + 1) for _ in range(5)]
    alfa, beta = 0.1, 3.0

    def run(a, b):
        y[0][0] = 0.0
        y[1][0] = 0.0
        if adiabatic:
            y[4][0] = 0.0
            y[2][0] = a
            y[3][0] = b
        else:
            y[3][0] = t_wall
            y[2][0] = a
            y[4][0] = b
        _integrate(n, h, y, c, ma, pr, gamma)
        return y[1][-1], y[3][-1]

    it = 0
    for it in range(1, max_iter + 1):
        y2o, y4o = run(alfa, beta)
        y2a, y4a = run(alfa + delta, beta)
        y2b, y4b = run(alfa, beta + delta)
        a11 = (y2a - y2o) / delta
        a21 = (y4a - y4o) / delta
        a12 = (y2b - y2o) / delta
        a22 = (y4b - y4o) / delta
        r1 = 1 - y2o
        r2 = 1 - y4o
        det = a11 * a22 - a12 * a21
        alfa = alfa + (a22 * r1 - a12 * r2) / det
        beta = beta + (a11 * r2 - a21 * r1) / det
        if abs(y[1][-1] - 1) < tol and abs(y[3][-1] - 1) < tol:
            break
    y4 = y[3]
    x = np.zeros(n + 1)
    for i in range(1, n + 1):
        x[i] = (eta[i] - 0) * (y4[0] + 2 * np.sum(y4[1:i]) + y4[i]) / (2 * eta[i]) * h
    return {"eta": x, "f": y[0].copy(), "f_p": y[1].copy(), "f_pp": y[2].copy(), "rbar": y[3].copy(),
            "rbar_p": y[4].copy(), "iterations": it}


def _makima(x, v, xq, fill):
    out = Akima1DInterpolator(x, v, method="makima")(xq)
    out = np.where((xq < x[0]) | (xq > x[-1]), fill, out)
    return out


def compressible_bl_flow(Re, Ma, Pr, T0, gamma, X, Y, Z, adiabatic=False, t_wall=None):
    """U, V, W, R, E of calcCompressibleBL.m:1-93 on the mesh X, Y, Z (nx x ny x nz each).

    ``t_wall``: the wall temperature in kelvin (flowParameters.Twall), default T0."""
    X, Y = np.asarray(X, float), np.asarray(Y, float)
    nz = len(np.atleast_1d(Z))
    xr = Re / BLASIUS_SCALE ** 2
    e0 = 1 / (gamma * (gamma - 1) * Ma ** 2)
    tw = 1.0 if t_wall is None else t_wall / T0
    sol = solve_compressible_bl(Ma, Pr, T0, tw, gamma, adiabatic)
    eta = sol["eta"]
    u_r = sol["f_p"]
    r_r = 1 / sol["rbar"]
    e_r = sol["rbar"] * e0
    v_r = -sol["rbar"] * (sol["f"] - eta * sol["f_p"]) * (BLASIUS_SCALE / np.sqrt(2)) * (1 / Re)
    y_r = eta * np.sqrt(2) / BLASIUS_SCALE
    nx, ny = len(X), len(Y)
    U = np.ones((nx, ny))
    V = np.zeros((nx, ny))
    W = np.zeros((nx, ny))
    R = np.ones((nx, ny))
    E = np.ones((nx, ny)) * e0
    iy = np.flatnonzero(Y >= 0)
    for ix in np.flatnonzero(X > 0):
        y_l = y_r * np.sqrt(X[ix] / xr)
        U[ix, iy] = _makima(y_l, u_r, Y[iy], u_r[-1])
        V[ix, iy] = _makima(y_l, v_r, Y[iy], v_r[-1])
        R[ix, iy] = _makima(y_l, r_r, Y[iy], r_r[-1])
        E[ix, iy] = _makima(y_l, e_r, Y[iy], e_r[-1])
    U[np.ix_(X == 0, Y == 0)] = 0
    below = np.flatnonzero(Y < 0)
    at0 = np.flatnonzero(Y == 0)
    if below.size and at0.size:
        for A in (U, V, R, E):
            A[:, below] = A[:, at0[:1]]
    rep = lambda A: np.asfortranarray(np.repeat(A[:, :, None], nz, axis=2))
    return {"U": rep(U), "V": rep(V), "W": rep(W), "R": rep(R), "E": rep(E), "solution": sol}
