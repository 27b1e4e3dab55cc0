"""The linearised boundary layer at its full size: 512 x 192 x 64, spanwise periodic, about
a steady base from an SFD run of the 2-D boundary layer replicated in z (the reference's
steady base of main.F90:381-402 and runDNSautomean.m), frozen SFD damping in the buffers.

* Dot-product test of one tangent and one adjoint step, <J v, w> = <v, J^T w> within 1e-11.
* Gradient of E(q') = 1/2 ||q'(T)||^2 over 200 steps, checkpointed every 100 steps (two
  recompute segments), against the central difference of E along a random direction: E is
  quadratic in q'(0), so the central difference with unit step is exact up to round-off
  (bar 1e-9 of |g|); and along q'(0) itself, where <g, q'(0)> = 2 E (Euler's identity for
  a quadratic form, bar 1e-11).
"""
import pytest

pytestmark = pytest.mark.gpu

NX, NY, NZ = 512, 192, 64


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd.adjoint import Propagator, state_tensor
    from gatt_dns_amd.baseflow import extrude, sfd_base
    from gatt_dns_amd.prep import cases
    print("2-D SFD base 512x192", flush=True)
    _, q2, info = sfd_base(NX, NY, steps=5000)
    print("base after", info, flush=True)
    case = cases.boundary_layer(nx=NX, ny=NY, nz=NZ, noise=0.0, sfd_delta=-1.0)
    base = extrude(q2, NZ)
    prop = Propagator(case)
    q0 = state_tensor(case, base)
    prop.load(q0)
    prop.solver.apply_bound_cond(0.0)
    q0 = prop.save()
    yield prop, q0
    prop.close()
    torch.cuda.empty_cache()


def _unit(prop, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(prop.shape, generator=g, dtype=torch.float64, device="cuda")
    return v / torch.linalg.norm(v)


def test_dot_product_one_step_full_size(setup):
    prop, q0 = setup
    v, w = _unit(prop, 5), _unit(prop, 6)
    print("tangent and adjoint step at 512x192x64", flush=True)
    _, jv = prop.tangent(q0, v, 1)
    jtw = prop.adjoint(q0, w)
    lhs = float((jv * w).sum())
    rhs = float((v * jtw).sum())
    scale = float(jv.norm() * w.norm())
    print("<Jv,w> %.17e <v,J^T w> %.17e" % (lhs, rhs), flush=True)
    assert abs(lhs - rhs) <= 1e-11 * scale


def test_gradient_200_steps_checkpointed(setup):
    from gatt_dns_amd.adjoint import optimal_energy_gradient
    prop, q0 = setup
    dq0, v = _unit(prop, 7), _unit(prop, 8)
    n = 200

    def energy(dq):
        prop.load(q0)
        prop.load_perturbation(dq)
        for s in range(n):
            prop.solver.tlm_step(prop.t0 + s * prop.dt, prop.dt)
        qT = prop.save_perturbation()
        return 0.5 * float((qT * qT).sum())

    print("gradient over 200 steps, checkpoints every 100", flush=True)
    e0, g = optimal_energy_gradient(prop, q0, dq0, n, checkpoint_every=100)
    assert e0 == pytest.approx(energy(dq0), rel=1e-13)
    assert float((g * dq0).sum()) == pytest.approx(2 * e0, rel=1e-11)
    fd = (energy(dq0 + v) - energy(dq0 - v)) / 2.0
    dg = float((g * v).sum())
    print("E %.6e, <g,v> %.12e, central difference %.12e" % (e0, dg, fd), flush=True)
    assert abs(fd - dg) <= 1e-9 * float(g.norm())
