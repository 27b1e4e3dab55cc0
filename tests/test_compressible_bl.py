"""The compressible boundary-layer initial flow (calcCompressibleBL.m), pinned by what
the similarity equations fix rather than by MATLAB (absent here):

* incompressible limit (Ma -> 0, Twall = 1): T = 1 and f''(0) is Blasius's 0.4696 of
  f''' + f f'' = 0;
* Pr = 1: the Crocco-Busemann integral T = Tw + (1 - Tw + (gamma-1) Ma^2 / 2) f'
  - (gamma-1) Ma^2 / 2 f'^2 holds for any viscosity law, for an isothermal wall and for the
  adiabatic wall, whose temperature is then the recovery value 1 + (gamma-1) Ma^2 / 2;
* Pr = 0.71, adiabatic: the laminar recovery factor (Tw - 1) / ((gamma-1) Ma^2 / 2) is
  close to sqrt(Pr);
* the boundary conditions at the wall and at the edge;
* the flow on a mesh: the free stream for x <= 0, U = 0 at the wall, the edge values far
  from it, and self-similarity in y / sqrt(x).
"""
import numpy as np
import pytest

from gatt_dns_amd.prep.compressible_bl import compressible_bl_flow, solve_compressible_bl


def test_incompressible_limit():
    s = solve_compressible_bl(1e-4, 0.71, 300.0, 1.0, 1.4, False)
    assert s["f_pp"][0] == pytest.approx(0.4696, abs=1e-4)
    assert np.max(np.abs(s["rbar"] - 1)) < 1e-8
    assert s["f"][0] == 0 and s["f_p"][0] == 0


@pytest.mark.parametrize("adiabatic,tw", [(False, 1.5), (False, 0.8), (True, 1.0)])
def test_crocco_busemann(adiabatic, tw):
    g, ma = 1.4, 0.5
    s = solve_compressible_bl(ma, 1.0, 300.0, tw, g, adiabatic)
    c = (g - 1) * ma ** 2 / 2
    t_wall = s["rbar"][0]
    cb = t_wall + (1 - t_wall + c) * s["f_p"] - c * s["f_p"] ** 2
    assert np.max(np.abs(s["rbar"] - cb)) < 1e-8
    if adiabatic:
        assert t_wall == pytest.approx(1 + c, abs=1e-8)
        assert s["rbar_p"][0] == 0
    else:
        assert t_wall == tw
    assert abs(s["f_p"][-1] - 1) < 1e-8 and abs(s["rbar"][-1] - 1) < 1e-8


def test_recovery_factor():
    s = solve_compressible_bl(0.5, 0.71, 300.0, 1.0, 1.4, True)
    r = (s["rbar"][0] - 1) / (0.4 * 0.25 / 2)
    assert abs(r - np.sqrt(0.71)) < 0.005


def test_flow_on_mesh():
    Re, Ma = 1000.0, 0.5
    X = np.linspace(-2.0, 40.0, 43)
    Y = np.concatenate([[-0.5], np.linspace(0.0, 12.0, 61)])
    fl = compressible_bl_flow(Re, Ma, 0.71, 300.0, 1.4, X, Y, np.zeros(2))
    e0 = 1 / (1.4 * 0.4 * Ma ** 2)
    U, V, R, E = fl["U"][..., 0], fl["V"][..., 0], fl["R"][..., 0], fl["E"][..., 0]
    assert fl["U"].shape == (43, 62, 2) and np.all(fl["W"] == 0)
    up = X <= 0
    np.testing.assert_array_equal(U[up & (X < 0)], 1.0)
    assert U[X == 0, 1] == 0.0                      # the leading-edge node
    wall = 1                                        # Y == 0
    assert np.all(U[X > 0, wall] == 0.0)
    np.testing.assert_allclose(E[X > 0, wall], e0, rtol=1e-12)   # isothermal wall at T0
    np.testing.assert_array_equal(U[:, 0], U[:, wall])          # the row below y = 0 copies y = 0
    # far from the wall: the edge values
    np.testing.assert_allclose(U[X > 0, -1], 1.0, atol=1e-6)
    np.testing.assert_allclose(R[X > 0, -1], 1.0, atol=1e-6)
    # self-similarity: U(x, y) = U(4x, 2y)
    fl2 = compressible_bl_flow(Re, Ma, 0.71, 300.0, 1.4, np.array([10.0, 40.0]), np.array([0.3, 0.6]), np.zeros(1))
    assert fl2["U"][0, 0, 0] == pytest.approx(fl2["U"][1, 1, 0], rel=1e-12)
    assert np.all(np.isfinite(V)) and np.all(V[X > 0, -1] > 0)  # displacement: V > 0 at the edge
