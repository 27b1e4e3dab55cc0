"""GPU parity on flow regions with walls inside the domain: the boundary layer over a cavity
cut into the plate with a roughness element downstream (boundaryLayerIsothermal.m:338-358, the
reference's parameters.m configuration in small) and the lid-driven cavity with a block inside
(lidDrivenFlow.m:16-25).  The walls bring line types split by wall segments, interior wall
faces in every orientation and convex and concave corners.

Each field against its own scale: the boundary conditions 1e-14, one step 1e-12, 10 steps
1e-10, with the SFD mean of the boundary layer at the same bars.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu

INF = np.inf


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


def _cavity_bl():
    kw = dict(nx=96, ny=48, nbuf=10, n_below=8, wall_depth=2.0)
    c0 = cases.boundary_layer(**kw)
    X, Y = np.asarray(c0.X), np.asarray(c0.Y)
    return cases.boundary_layer(cavities=[dict(x=(X[30], X[40]), y=(-2.0, 0.0), z=(-INF, INF))],
                                roughness=[dict(x=(X[50], X[53]), y=(0.0, Y[12]), z=(-INF, INF))], **kw)


def _lid_block():
    return cases.lid_driven(nx=48, ny=40, roughness=[dict(x=(0.35, 0.6), y=(0.3, 0.55), z=(-INF, INF))])


CASES = {"cavity_bl": _cavity_bl, "lid_block": _lid_block}
FIELDS = "UVRE"


@pytest.mark.parametrize("name", list(CASES))
def test_boundary_conditions(rt, oracle_mod, name):
    case = CASES[name]()
    print(name, "boundary conditions", flush=True)
    rng = np.random.default_rng(3)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in case.state0.items()}
    q["W"][:] = 0.0
    want = oracle_mod.Oracle(case).apply_bound_cond(q, 0.0)
    s = rt.Solver(case)
    s.set_state(q)
    s.apply_bound_cond(0.0)
    got = s.get_state()
    s.close()
    for k in FIELDS:
        assert rel_err(got[k], want[k]) <= 1e-14, (name, k, rel_err(got[k], want[k]))


@pytest.mark.parametrize("name", list(CASES))
def test_line_solves(rt, oracle_mod, name):
    import torch
    case = CASES[name]()
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
            assert rel_err(got, want) < 1e-13, (name, d, kind, rel_err(got, want))
    s.close()


@pytest.mark.parametrize("name", list(CASES))
def test_steps(rt, oracle_mod, name):
    case = CASES[name]()
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    if case.sfd_type > 0:
        o.init_mean(q0)
    s = rt.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    want, t = q0, 0.0
    for nsteps, tol in ((1, 1e-12), (9, 1e-10)):
        print(name, "steps", nsteps, flush=True)
        want, _, dv_w = o.run(want, t, case.dt, nsteps)
        _, dv = s.advance(nsteps, t, case.dt)
        t += nsteps * case.dt
        got = s.get_state()
        assert dv == dv_w
        errs = {k: rel_err(got[k], want[k]) for k in FIELDS}
        print(name, nsteps, errs, flush=True)
        for k in FIELDS:
            assert errs[k] <= tol, (name, nsteps, k, errs[k])
        if case.sfd_type > 0:
            gm, wm = s.get_mean(), o.mean()
            for k in FIELDS:
                assert rel_err(gm[k], wm[k]) <= tol, (name, "mean", k)
    s.close()


TRIM_CASES = {
    "bl160": lambda: cases.boundary_layer(nx=160, ny=64, nbuf=12),
    "adiabatic": lambda: cases.boundary_layer(nx=160, ny=64, nbuf=12, adiabatic=True),
    "pressure_inlet": lambda: cases.boundary_layer(nx=160, ny=64, nbuf=12, pressure_inlet=True),
    "free_slip": lambda: cases.boundary_layer(nx=160, ny=64, nbuf=12, free_slip=True),
    "cavity_bl": _cavity_bl,
    "bl3d": lambda: cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12),
    "lid_block": _lid_block,
}


@pytest.mark.parametrize("name", list(TRIM_CASES))
def test_boundary_program_trimmed_equals_listed(rt, name, monkeypatch):
    """The boundary program without its dead writes (an op's points that a later op overwrites
    unread; the default where that takes fewer waves) gives the states of the program as
    listed bit for bit: one application on a perturbed state, then 3 steps."""
    case = TRIM_CASES[name]()
    rng = np.random.default_rng(8)
    q = {k: np.asfortranarray(v + 0.01 * rng.standard_normal(v.shape)) for k, v in case.state0.items()}
    if case.nz == 1:
        q["W"][:] = 0.0
    out = []
    for trim in ("1", "0"):
        monkeypatch.setenv("GATT_BC_TRIM", trim)
        s = rt.Solver(case)
        s.set_state(q)
        s.apply_bound_cond(0.0)
        once = s.get_state()
        s.advance(3, 0.0, case.dt)
        out.append((once, s.get_state()))
        s.close()
    for k in "UVWRE":
        assert np.array_equal(out[0][0][k], out[1][0][k]), (name, "application", k)
        assert np.array_equal(out[0][1][k], out[1][1][k]), (name, "steps", k)
