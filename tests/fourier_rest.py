"""Closed-form spectrum of the linearised time step about fluid at rest in a periodic box.

About U = V = W = 0, R = 1, E = E0 on a uniform periodic mesh every operator
of the step is a circulant, so Fourier modes decouple.  With the compact
first-derivative symbol (centred LHS ``l``, RHS ``r``, h = 1/n)

    d(theta) = i n (2 sum_k r_k sin k theta) / (l_0 + 2 sum_k l_k cos k theta)

and the filter transfer function

    T(theta) = (r_0 + 2 sum_k r_k cos k theta) / (1 + 2 alpha_f cos theta)

the linearised Navier-Stokes RHS (nseq.F90:218-318 with zero base gradients:
stresses and heat flux are linear in the perturbation, mu = mu(E0)) is, per
mode, a small matrix L acting on (u, v, [w], rho, e):

    div  = d_x u + d_y v [+ d_z w]
    du_a = sum_b d_b tau_ab - (gamma-1)(d_a e + E0 d_a rho)
    drho = -div
    de   = -P0 div + gamma mu/(Re Pr) (d_x^2 + d_y^2 [+ d_z^2]) e
    tau_ab = mu/Re (d_b u_a + d_a u_b - 2/3 delta_ab div)

(second derivatives are the first-derivative operator applied twice, as in the
reference).  One step is RK4 (timeIntegration.F90) followed by the filter in
each direction, so its amplification is ``T_x T_y [T_z] * sum_{j<=4} (dt L)^j/j!``
and the time-T propagator is that matrix to the power nsteps.

Test infrastructure only.
"""
import itertools

import numpy as np

from gatt_dns_amd.prep.stencils import finite_difference_coefficients, spatial_filter_coefficients


def derivative_symbol(n, method="SL4"):
    cl, cr, _, _ = finite_difference_coefficients(method)
    th = 2 * np.pi * np.arange(n) / n
    num = sum(2j * cr[k] * np.sin(k * th) for k in range(1, len(cr)))
    den = cl[0] + sum(2 * cl[k] * np.cos(k * th) for k in range(1, len(cl)))
    return n * num / den


def filter_symbol(n, alpha_f=0.49, borders="reducedOrder"):
    fl, fr, _, _ = spatial_filter_coefficients(alpha_f, borders)
    th = 2 * np.pi * np.arange(n) / n
    num = fr[0] + sum(2 * fr[k] * np.cos(k * th) for k in range(1, len(fr)))
    den = fl[0] + sum(2 * fl[k] * np.cos(k * th) for k in range(1, len(fl)))
    return num / den


def mode_operator(dvec, case):
    """L for one mode; dvec = (d_x, d_y[, d_z]) symbols.  Slots u, v, [w], rho, e."""
    g, Re, Pr, Ma, T0 = case.gamma, case.Re, case.Pr, case.Ma, case.T0
    E0 = case.state0["E"].flat[0]
    mu2 = g * (g - 1) * Ma * Ma * E0
    mu = (1 + 110.0 / T0) / (mu2 + 110.0 / T0) * mu2 * np.sqrt(mu2)
    nd = len(dvec)
    ns = nd + 2
    ir, ie = nd, nd + 1
    L = np.zeros((ns, ns), dtype=complex)
    c = mu / Re
    # div as a row over slots
    div = np.zeros(ns, dtype=complex)
    div[:nd] = dvec
    for a in range(nd):
        row = np.zeros(ns, dtype=complex)
        for b in range(nd):
            tau = np.zeros(ns, dtype=complex)      # tau_ab as a row
            tau[a] += c * dvec[b]
            tau[b] += c * dvec[a]
            if a == b:
                tau -= c * 2.0 / 3.0 * div
            row += dvec[b] * tau
        row[ie] -= (g - 1) * dvec[a]
        row[ir] -= (g - 1) * E0 * dvec[a]
        L[a] = row
    L[ir] = -div
    P0 = (g - 1) * E0
    L[ie] = -P0 * div
    L[ie, ie] += g * mu / (Re * Pr) * sum(d * d for d in dvec)
    return L


def rk4_amplification(L, dt):
    A = dt * L
    G = np.eye(L.shape[0], dtype=complex)
    term = np.eye(L.shape[0], dtype=complex)
    for j in range(1, 5):
        term = term @ A / j
        G = G + term
    return G


def spectral_z_symbol(nz):
    """Symbol of the 4-point spectral z stencil (LHS 1, RHS [0, pi/4], makeMatrices.m:28-38) at
    the first z harmonic on a z period of 1: i nz 2 (pi/4) sin(2 pi/nz) = 2 pi i for nz = 4."""
    return 1j * nz * 2 * (np.pi / 4) * np.sin(2 * np.pi / nz)


def propagator_eigenvalues(case, nsteps, dt, single_beta=False):
    """All eigenvalues mu of the time-(nsteps*dt) propagator (one per mode and slot).

    With ``single_beta`` the modes are the x-y modes times the first z harmonic
    (U,V,R,E ~ cos, W ~ sin: the subspace the singleBeta packing spans); the z
    filter of a 4-point z mesh is the identity.
    """
    dims = [case.nx, case.ny] + ([case.nz] if case.nz > 1 and not single_beta else [])
    dsym = [derivative_symbol(n) for n in dims]
    fsym = [filter_symbol(n) for n in dims]
    out = []
    for idx in itertools.product(*[range(n) for n in dims]):
        dvec = [dsym[k][i] for k, i in enumerate(idx)]
        if single_beta:
            dvec.append(spectral_z_symbol(case.nz))
        F = np.prod([fsym[k][i] for k, i in enumerate(idx)])
        A = F * rk4_amplification(mode_operator(dvec, case), dt)
        out.extend(np.linalg.eigvals(A) ** nsteps)
    return np.array(out)


def match(mu_a, mu_b):
    """Largest distance of an optimal one-to-one matching of two eigenvalue sets."""
    from scipy.optimize import linear_sum_assignment
    cost = np.abs(mu_a[:, None] - mu_b[None, :])
    r, c = linear_sum_assignment(cost)
    return cost[r, c].max(), r, c
