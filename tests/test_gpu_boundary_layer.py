"""GPU parity on the isothermal boundary layer (boundaryLayerIsothermal.m): stretched mesh, EX4
buffers, decentred filter closures switched off at the x ends and the top, a free-slip region ahead
of the plate with its x break point, SFD type 2 (chi in the buffers, Delta = 10).

Tolerances as the other parity tests: line solves and boundary conditions 1e-13, one step 1e-12,
20 steps 1e-10; the 1024 x 256 configuration one step 1e-12 and 3 steps 1e-10.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu

FIELDS = "UVRE"


def rel_err(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale else float(np.max(np.abs(a)))


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd import runtime
    runtime.load_library()
    return runtime


@pytest.fixture(scope="module")
def small():
    return cases.boundary_layer(nx=160, ny=64, nbuf=12)


def test_line_solves(rt, oracle_mod, small):
    import torch
    case = small
    o = oracle_mod.Oracle(case)
    s = rt.Solver(case)
    rng = np.random.default_rng(5)
    shape = (case.nx, case.ny, 1)
    for d in (0, 1):
        for kind in (0, 1):
            F = np.asfortranarray(rng.standard_normal(shape))
            want = o.solve(d, kind, F)
            tin = torch.from_numpy(F.ravel(order="F").copy()).cuda()
            tout = torch.zeros_like(tin)
            s.line_solve(d, kind, [tin], [tout])
            got = tout.cpu().numpy().reshape(shape, order="F")
            assert rel_err(got, want) < 1e-13, (d, kind, rel_err(got, want))


def test_boundary_conditions(rt, oracle_mod, small):
    case = small
    rng = np.random.default_rng(3)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in case.state0.items()}
    q["W"][:] = 0.0
    want = oracle_mod.Oracle(case).apply_bound_cond(q, 0.0)
    s = rt.Solver(case)
    s.set_state(q)
    s.apply_bound_cond(0.0)
    got = s.get_state()
    for k in FIELDS:
        assert rel_err(got[k], want[k]) <= 1e-14, (k, rel_err(got[k], want[k]))


@pytest.mark.parametrize("nsteps,tol", [(1, 1e-12), (20, 1e-10)])
def test_steps_small(rt, oracle_mod, small, nsteps, tol):
    case = small
    print("boundary layer 160x64, steps", nsteps, flush=True)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    want, mc_w, dv_w = o.run(q0, 0.0, case.dt, nsteps)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    mc, dv = s.advance(nsteps, 0.0, case.dt)
    got = s.get_state()
    for k in FIELDS:
        assert rel_err(got[k], want[k]) <= tol, (k, rel_err(got[k], want[k]))
    gm, wm = s.get_mean(), o.mean()
    for k in FIELDS:
        assert rel_err(gm[k], wm[k]) <= tol, k
    assert dv == dv_w


@pytest.mark.parametrize("variant", ["adiabatic", "pressure_inlet", "free_slip"])
def test_wall_variant_steps(rt, oracle_mod, variant):
    """boundaryLayerAdiabatic.m walls (zero-gradient E, adiabatic corners), the pressure
    conditions of boundaryLayerIsothermalPressureInlet.m and the slip plate of
    boundaryLayerFreeSlip.m: 10 steps."""
    case = cases.boundary_layer(nx=160, ny=64, nbuf=12, **{variant: True})
    print("boundary layer 160x64", variant, flush=True)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    want, _, dv_w = o.run(q0, 0.0, case.dt, 10)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    mc, dv = s.advance(10, 0.0, case.dt)
    got = s.get_state()
    assert dv == dv_w
    for k in FIELDS:
        assert rel_err(got[k], want[k]) <= 1e-10, (k, rel_err(got[k], want[k]))


def test_z_symmetry_steps(rt, oracle_mod):
    """boundaryLayerIsothermalSymmZ.m: symmetry planes at both z ends (non-periodic z). 4 steps."""
    case = cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12, z_symmetry=True)
    print("symmetric-z boundary layer 64x40x12", flush=True)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    want, _, dv_w = o.run(q0, 0.0, case.dt, 4)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    mc, dv = s.advance(4, 0.0, case.dt)
    got = s.get_state()
    assert dv == dv_w
    # every field against its own scale; W grows from the seeded 1e-5 noise, so its bar allows
    # for the round-off of the O(1) terms it is computed from
    errs = {k: rel_err(got[k], want[k]) for k in "UVWRE"}
    print("symmetric-z boundary layer", errs, flush=True)
    tol = {"U": 1e-10, "V": 1e-10, "W": 1e-8, "R": 1e-10, "E": 1e-10}
    for k, e in errs.items():
        assert e <= tol[k], (k, e)


# V is O(1e-3) here: after 3 steps its largest difference (2e-14 absolute, in the first row above
# the wall) is 3e-11 of max|V| while U, R and E stay at 1e-13 of theirs
@pytest.mark.parametrize("nsteps,tol", [(1, 1e-12), (3, 1e-10)])
def test_steps_1024x256(rt, oracle_mod, nsteps, tol):
    case = cases.boundary_layer()
    print("boundary layer 1024x256, steps", nsteps, flush=True)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    want, _, _ = o.run(q0, 0.0, case.dt, nsteps)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    s.advance(nsteps, 0.0, case.dt)
    got = s.get_state()
    errs = {k: rel_err(got[k], want[k]) for k in FIELDS}
    print("1024x256", nsteps, errs, flush=True)
    for k in FIELDS:
        assert errs[k] <= tol, (k, errs[k])


def test_spanwise_periodic_step(rt, oracle_mod):
    """The 3-D (z-periodic) extrusion: one step and five steps against the oracle."""
    case = cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    for nsteps, tol in ((1, 1e-12), (4, 1e-10)):
        want, _, _ = o.run(q0 if nsteps == 1 else want, 0.0 if nsteps == 1 else case.dt, case.dt, nsteps)
        s.advance(nsteps, 0.0 if nsteps == 1 else case.dt, case.dt)
        got = s.get_state()
        # every field against its own scale; W grows from the seeded 1e-5 noise, so its bar
        # allows for the round-off of the O(1) terms it is computed from
        errs = {k: rel_err(got[k], want[k]) for k in "UVWRE"}
        print("3-D boundary layer", nsteps, errs, flush=True)
        for k, e in errs.items():
            assert e <= (tol * 100 if k == "W" else tol), (nsteps, k, e)
