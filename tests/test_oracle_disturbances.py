"""Pins of the oracle's pressureGradient forcing and time-dependent disturbances.

* pressureGradient (pressureGradient.F90): the laminar channel parabola
  u = dpdx Re / (2 mu) (H^2/4 - y^2) is a steady solution of the x-momentum
  equation with the forcing dU += dpdx/R; on a uniform y mesh the compact
  derivatives of the quadratic are exact, so the discrete dU vanishes to
  rounding at every density.
* periodicSine (periodicSine.F90) and packet_2d (packet_2d.F90) against their
  closed forms in double precision; the routines' single-precision pi, alpha,
  x0, Tend, Tp, sigmaT and sigmaX leave a relative difference of about 1e-7.
  packet_2d multiplies the whole field value by the amplitude and does nothing
  once t > Tend.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases


@pytest.mark.parametrize("nz,rho", [(1, 1.0), (1, 2.0), (6, 2.0)])
def test_pressure_gradient_keeps_the_laminar_profile(oracle_mod, nz, rho):
    case = cases.poiseuille(nx=16, ny=33, nz=nz, y_mesh="uniform", spatial="SL4")
    q = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
    q["R"][:] = rho
    q["E"][:] = case.E0     # P = (gamma-1) rho E is uniform, T = 1
    o = oracle_mod.Oracle(case)
    d = o.rhs(q)
    dpdx = case.forcing[0].par[0]
    inner = d["U"][:, 1:-1, :]
    assert np.max(np.abs(inner)) <= 1e-10 * dpdx, np.max(np.abs(inner))
    # without the forcing the same state accelerates at -dpdx/R
    case.forcing = []
    d0 = oracle_mod.Oracle(case).rhs(q)
    assert np.allclose(d0["U"][:, 1:-1, :], -dpdx / rho, rtol=1e-8, atol=0)


def _wall_case(kind, par, yr=(-0.5, -0.5)):
    return cases.poiseuille(nx=32, ny=17, y_mesh="uniform", spatial="SL4",
                            disturb=[(kind, (1.0, 3.0), yr, "V", par)])


def test_periodic_sine_closed_form(oracle_mod):
    omega, amp = 3.0, 1e-3
    case = _wall_case("periodicSine", [omega, amp])
    ds = case.disturb[0]
    o = oracle_mod.Oracle(case)
    x = ds.X
    for t in (0.1, 0.37, 1.3):
        q = o.apply_bound_cond(case.state0, t)
        got = q["V"][ds.box[0] - 1:ds.box[1], 0, 0]
        want = amp * np.sin(2 * np.pi / (x[-1] - x[0]) * (x - x[0])) * np.sin(omega * t)
        assert np.max(np.abs(got - want)) <= 1e-6 * amp, t
        # nothing outside the box
        assert np.all(q["V"][ds.box[1]:, 0, 0] == 0.0)


@pytest.mark.parametrize("yr,v0", [((-0.5, -0.5), 0.0), ((0.0, 0.1), 0.25)])
def test_packet_2d_closed_form(oracle_mod, yr, v0):
    omega, nmodes, amp = 2.0, 5.0, 1e-3
    case = _wall_case("packet_2d", [omega, nmodes, amp], yr)
    ds = case.disturb[0]
    b = ds.box
    q0 = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
    q0["V"][b[0] - 1:b[1], b[2] - 1:b[3], :] = v0
    o = oracle_mod.Oracle(case)
    x = ds.X
    Tend = (2 * np.pi / omega) / 4
    Tp, sT, sX, x0 = Tend / 2, Tend / 4, (x[-1] - x[0]) / 8, (x[-1] + x[0]) / 2
    for t in (0.05, 0.4, 0.7):
        q = o.apply_bound_cond(q0, t)
        got = q["V"][b[0] - 1:b[1], b[2] - 1, 0]
        modes = sum(np.cos(m * omega * (t - Tp)) for m in range(1, int(nmodes) + 1))
        packet = np.exp(-(t - Tp) ** 2 / (2 * sT ** 2)) * np.exp(-(x - x0) ** 2 / (2 * sX ** 2)) * modes / nmodes
        base = 0.0 if yr[0] == -0.5 else v0     # the wall's Dirichlet V = 0 acts first
        want = amp * (base + packet)
        assert np.max(np.abs(got - want)) <= 1e-6 * max(amp, abs(amp * base)), (t, np.max(np.abs(got - want)))
    # after Tend the routine returns without touching the field
    q = o.apply_bound_cond(q0, Tend * 1.01)
    base = 0.0 if yr[0] == -0.5 else v0
    assert np.all(q["V"][b[0] - 1:b[1], b[2] - 1, 0] == base)
