"""GPU checks of the tangent-linear and adjoint steps.

* dot-product test <J v, w> = <v, J^T w> within 1e-11 (one step and a
  checkpointed multi-step propagator): 2-D and 3-D periodic boxes (3-D also with the
  turbulenceGenerator forcing, whose global sums the linearised steps differentiate), the lid-driven
  cavity (walls, corners, lid) and the boundary layer (inflow/outflow/top boxes,
  free-slip segment, buffers, SFD with a frozen mean, and with an evolving mean, whose
  five fields then belong to the state, the perturbation and the adjoint), and the boundary
  layer over a cavity with a roughness element;
* the tangent against central differences of the oracle's step (the
  reference's non-linear step restated on the CPU);
* the transposed line solves against the transposes of the forward solves
  (dense operators assembled column by column on the GPU);
* the gradient of E(T) = 1/2 ||q(T)||^2 from torch.autograd against a
  finite-difference directional derivative.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _case(kind):
    if kind == "box2d":
        c = cases.periodic_rest(32, 24)
        rng = np.random.Generator(np.random.PCG64(1))
        x = np.arange(32) / 32
        y = np.arange(24) / 24
        X, Y = np.meshgrid(x, y, indexing="ij")
        q = {k: np.array(v, copy=True) for k, v in c.state0.items()}
        q["U"] = q["U"] + (0.3 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal((32, 24)))[:, :, None]
        q["V"] = q["V"] + (0.2 * np.cos(2 * np.pi * X))[:, :, None]
        q["R"] = q["R"] + 0.05 * np.cos(2 * np.pi * (X + Y))[:, :, None]
        q["E"] = q["E"] * (1 + 0.02 * np.sin(4 * np.pi * X))[:, :, None]
        return c, q
    if kind == "cavity":
        # walls: Dirichlet and pressure Neumann boxes, corners, the lid profile
        c = cases.lid_driven(nx=29, ny=23)
        rng = np.random.Generator(np.random.PCG64(2))
        q = {k: np.array(v, copy=True) for k, v in c.state0.items()}
        x = np.linspace(0, 1, 29)[:, None, None]
        y = np.linspace(0, 1, 23)[None, :, None]
        q["U"] = q["U"] + 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y) + 0.01 * rng.standard_normal(q["U"].shape)
        q["V"] = q["V"] + 0.1 * np.sin(2 * np.pi * x) * np.sin(np.pi * y)
        q["R"] = q["R"] + 0.02 * np.cos(np.pi * x)
        q["E"] = q["E"] * (1 + 0.01 * np.cos(np.pi * y))
        return c, {k: np.asfortranarray(v) for k, v in q.items()}
    if kind == "bl3d":
        # the adjoint configuration in small: spanwise-periodic boundary layer with walls
        c = cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12, sfd_delta=-1.0)
        q = {k: np.array(v, copy=True, order="F") for k, v in c.state0.items()}
        z = np.arange(12) / 12
        q["W"] = np.asfortranarray(q["W"] + 1e-3 * np.sin(2 * np.pi * z)[None, None, :] * c.flow_region)
        return c, q
    if kind == "bl_cavity":
        # a cavity cut into the plate and a roughness element on it: interior wall faces in
        # every orientation, convex and concave corners, line types split by walls
        kw = dict(nx=96, ny=48, nbuf=10, n_below=8, wall_depth=2.0, sfd_delta=-1.0)
        c0 = cases.boundary_layer(**kw)
        X, Y = np.asarray(c0.X), np.asarray(c0.Y)
        c = cases.boundary_layer(cavities=[dict(x=(X[30], X[40]), y=(-2.0, 0.0), z=(-np.inf, np.inf))],
                                 roughness=[dict(x=(X[50], X[53]), y=(0.0, Y[12]), z=(-np.inf, np.inf))], **kw)
        return c, {k: np.array(v, copy=True, order="F") for k, v in c.state0.items()}
    if kind == "bl":
        # inflow, outflow and top boxes (null second derivatives), the wall with its free-slip
        # segment, EX4 buffers, SFD with a frozen mean in the buffers
        c = cases.boundary_layer(nx=96, ny=48, nbuf=10, sfd_delta=-1.0)
        return c, {k: np.array(v, copy=True, order="F") for k, v in c.state0.items()}
    # box3d: at rest of any forcing; box3d_forced: with turbulenceGenerator (global sums)
    c = cases.turbulence_box(16, seed=3, forcing=kind == "box3d_forced")
    q = {k: np.array(v, copy=True) for k, v in c.state0.items()}
    q["R"] = q["R"] + 0.03 * np.sin(2 * np.pi * np.arange(16) / 16)[None, :, None]
    return c, q


def _rand(prop, seed):
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    v = rng.standard_normal(prop.shape)
    if prop.case.nz == 1:
        v[2::5] = 0.0     # W (and the mean's W)
    return torch.from_numpy(v).cuda()


SFD_KINDS = ["bl_sfd", "bl_sfd1", "bl3d_sfd"]


def _sfd_case(kind):
    """Boundary layers whose SFD mean evolves (Delta > 0): chi in the buffers (type 2) or
    everywhere (type 1), a strong blend (chi = 0.5, Delta = 0.05, so dt/Delta is about 0.5 and
    dt chi about 0.01) so that the coupling of state and mean weighs in every check, and a
    mean that starts away from the state."""
    nz = 12 if kind == "bl3d_sfd" else 1
    kw = dict(nx=64, ny=40, nbuf=8, nz=nz) if nz > 1 else dict(nx=96, ny=48, nbuf=10)
    c = cases.boundary_layer(sfd_chi=0.5, sfd_delta=0.05, sfd_type=1 if kind == "bl_sfd1" else 2, **kw)
    q = {k: np.array(v, copy=True, order="F") for k, v in c.state0.items()}
    rng = np.random.Generator(np.random.PCG64(23))
    if nz > 1:
        z = np.arange(nz) / nz
        q["W"] = np.asfortranarray(q["W"] + 1e-3 * np.sin(2 * np.pi * z)[None, None, :] * c.flow_region)
    mean = {k: np.asfortranarray(v * (1.0 + 1e-3 * rng.standard_normal(v.shape))) for k, v in q.items()}
    if nz > 1:
        mean["W"] = np.asfortranarray(q["W"] + 1e-4 * rng.standard_normal(q["W"].shape))
    return c, q, mean


@pytest.mark.parametrize("kind", SFD_KINDS)
def test_evolving_mean_dot_product_one_step(torch_cuda, kind):
    """<J (v, mu), (w, nu)> = <(v, mu), J^T (w, nu)> with the SFD mean in the state; the adjoint
    step leaves the base state and the base mean as given."""
    import torch
    from gatt_dns_amd.adjoint import Propagator
    case, q, mean = _sfd_case(kind)
    prop = Propagator(case)
    assert prop.shape[0] == 10
    q0 = prop.state_tensor(q, mean)
    for seed in (1, 2):
        v, w = _rand(prop, seed), _rand(prop, seed + 10)
        _, Jv = prop.tangent(q0, v)
        JTw = prop.adjoint(q0, w)
        assert torch.equal(prop.save(), q0)
        a = float((Jv * w).sum())
        b = float((v * JTw).sum())
        assert abs(a - b) <= 1e-11 * max(abs(a), abs(b)), (a, b, abs(a - b) / abs(a))
        # the mean's rows of the adjoint carry the coupling (not the identity)
        assert float((JTw[5:] - w[5:]).abs().max()) > 1e-3
    prop.close()


@pytest.mark.parametrize("kind", SFD_KINDS)
def test_evolving_mean_tangent_matches_oracle_differences(torch_cuda, kind, oracle_mod):
    """The tangent of (state, mean) against central differences of the oracle's step, whose
    SFD block updates the mean before the blend (main.F90:382-401)."""
    from gatt_dns_amd.adjoint import Propagator
    case, q, mean = _sfd_case(kind)
    prop = Propagator(case)
    q0 = prop.state_tensor(q, mean)
    v = _rand(prop, 5)
    _, Jv = prop.tangent(q0, v)
    Jv = Jv.cpu().numpy()
    vh = v.cpu().numpy()
    eps = 1e-5
    outs = []
    for sgn in (1.0, -1.0):
        o = oracle_mod.Oracle(case)
        qq = {k: np.asfortranarray(q[k] + sgn * eps * np.transpose(vh[i], (2, 1, 0))) for i, k in enumerate("UVWRE")}
        mm = {k: np.asfortranarray(mean[k] + sgn * eps * np.transpose(vh[5 + i], (2, 1, 0)))
              for i, k in enumerate("UVWRE")}
        o.init_mean(mm)
        st, _, _, _ = o.step(qq, 0.0, case.dt)
        mo = o.mean()
        outs.append(np.stack([np.transpose(st[k], (2, 1, 0)) for k in "UVWRE"]
                             + [np.transpose(mo[k], (2, 1, 0)) for k in "UVWRE"]))
    fd = (outs[0] - outs[1]) / (2 * eps)
    if case.nz == 1:
        fd[2::5] = 0.0
    for part in (slice(0, 5), slice(5, 10)):
        rel = np.linalg.norm(Jv[part] - fd[part]) / np.linalg.norm(fd[part])
        assert rel < 1e-7, (part, rel)
    prop.close()


@pytest.mark.parametrize("kind", ["bl_sfd", "bl3d_sfd"])
def test_evolving_mean_checkpointed_gradient(torch_cuda, kind):
    """torch.autograd through checkpointed steps of (state, mean) against finite differences, and
    <lam, v> = ||J^n v||^2 through the tangent and the adjoint."""
    import torch
    from gatt_dns_amd.adjoint import Propagator, advance, optimal_energy_gradient
    case, q, mean = _sfd_case(kind)
    prop = Propagator(case)
    q0 = prop.state_tensor(q, mean)
    nsteps, every = 5, 2
    qref = q0.clone()
    x = q0.clone().requires_grad_(True)
    J = 0.5 * ((advance(prop, x, nsteps, checkpoint_every=every) - qref) ** 2).sum()
    J.backward()
    g = x.grad.detach()
    v = _rand(prop, 9) * 1e-3
    eps = 1e-3

    def Jof(z):
        with torch.no_grad():
            return float(0.5 * ((advance(prop, z, nsteps, checkpoint_every=every) - qref) ** 2).sum())

    fd = (Jof(q0 + eps * v) - Jof(q0 - eps * v)) / (2 * eps)
    an = float((g * v).sum())
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
    _, Jv = prop.tangent(q0, v, n=nsteps)
    E, lam = optimal_energy_gradient(prop, q0, v, nsteps, checkpoint_every=every)
    assert abs(E - 0.5 * float((Jv * Jv).sum())) <= 1e-12 * E
    a = float((lam * v).sum())
    b = float((Jv * Jv).sum())
    assert abs(a - b) <= 1e-11 * abs(b), (a, b)
    prop.close()


@pytest.mark.parametrize("kind", ["box2d", "box3d", "box3d_forced", "cavity", "bl", "bl3d", "bl_cavity"])
def test_dot_product_one_step(torch_cuda, kind):
    from gatt_dns_amd.adjoint import Propagator, state_tensor
    case, q = _case(kind)
    prop = Propagator(case)
    q0 = state_tensor(case, q)
    for seed in (1, 2):
        v, w = _rand(prop, seed), _rand(prop, seed + 10)
        _, Jv = prop.tangent(q0, v)
        JTw = prop.adjoint(q0, w)
        a = float((Jv * w).sum())
        b = float((v * JTw).sum())
        assert abs(a - b) <= 1e-11 * max(abs(a), abs(b)), (a, b, abs(a - b) / abs(a))
    prop.close()


@pytest.mark.parametrize("kind", ["box2d", "box3d", "box3d_forced", "cavity", "bl", "bl3d", "bl_cavity"])
def test_tangent_matches_oracle_differences(torch_cuda, kind, oracle_mod):
    from gatt_dns_amd.adjoint import Propagator, state_tensor
    case, q = _case(kind)
    prop = Propagator(case)
    q0 = state_tensor(case, q)
    v = _rand(prop, 5)
    _, Jv = prop.tangent(q0, v)
    Jv = Jv.cpu().numpy()
    o = oracle_mod.Oracle(case)
    vh = v.cpu().numpy()
    eps = 1e-5
    outs = []
    for sgn in (1.0, -1.0):
        qq = {k: np.asfortranarray(q[k] + sgn * eps * np.transpose(vh[i], (2, 1, 0)))
              for i, k in enumerate("UVWRE")}
        st, _, _, _ = o.step(qq, 0.0, case.dt)
        outs.append(np.stack([np.transpose(st[k], (2, 1, 0)) for k in "UVWRE"]))
    fd = (outs[0] - outs[1]) / (2 * eps)
    if case.nz == 1:
        fd[2] = 0.0
    rel = np.linalg.norm(Jv - fd) / np.linalg.norm(fd)
    assert rel < 1e-7, rel
    prop.close()


@pytest.mark.parametrize("name", ["periodic", "cavity"])
def test_transposed_solves_are_transposes(torch_cuda, name):
    """Dense operators: gatt_line_solve_t equals the transpose of gatt_line_solve (deriv, filter; x, y)."""
    from gatt_dns_amd.runtime import Solver
    torch = torch_cuda
    case = cases.periodic_rest(12, 14) if name == "periodic" else cases.lid_driven(nx=15, ny=13)
    s = Solver(case)
    n = case.nx * case.ny
    for d in (0, 1):
        for kind in (0, 1):
            A = np.zeros((n, n))
            AT = np.zeros((n, n))
            for j in range(n):
                e = torch.zeros(n, dtype=torch.float64, device="cuda")
                e[j] = 1.0
                out = torch.zeros_like(e)
                s.line_solve(d, kind, [e.clone()], [out])
                A[:, j] = out.cpu().numpy()
                out_t = torch.zeros_like(e)
                s.line_solve_t(d, kind, [e.clone()], [out_t])
                AT[:, j] = out_t.cpu().numpy()
            assert np.abs(A).max() > 0.1
            np.testing.assert_allclose(AT, A.T, rtol=0, atol=1e-12 * np.abs(A).max())
    s.close()


@pytest.mark.parametrize("kind", ["box2d", "box3d", "box3d_forced", "cavity", "bl"])
def test_checkpointed_propagator_dot_product_and_gradient(torch_cuda, kind):
    import torch
    from gatt_dns_amd.adjoint import Propagator, advance, optimal_energy_gradient, state_tensor
    case, q = _case(kind)
    prop = Propagator(case)
    q0 = state_tensor(case, q)
    nsteps, every = 7, 3
    # J = 1/2 ||q(T) - q_ref||^2 through torch.autograd
    qref = q0.clone()
    x = q0.clone().requires_grad_(True)
    qT = advance(prop, x, nsteps, checkpoint_every=every)
    J = 0.5 * ((qT - qref) ** 2).sum()
    J.backward()
    g = x.grad.detach()
    v = _rand(prop, 9) * 1e-3
    eps = 1e-3        # perturbations of 1e-6: truncation ~1e-12 relative, rounding noise ~1e-9

    def Jof(z):
        with torch.no_grad():
            out = advance(prop, z, nsteps, checkpoint_every=every)
            return float(0.5 * ((out - qref) ** 2).sum())

    fd = (Jof(q0 + eps * v) - Jof(q0 - eps * v)) / (2 * eps)
    an = float((g * v).sum())
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
    # the linearised propagator: <J^n v, w> = <v, (J^n)^T w> through the tangent and the adjoint
    w = _rand(prop, 11)
    _, Jv = prop.tangent(q0, v, n=nsteps)
    E, lam = optimal_energy_gradient(prop, q0, v, nsteps, checkpoint_every=every)
    assert abs(E - 0.5 * float((Jv * Jv).sum())) <= 1e-12 * E
    # d E / d v = (J^n)^T J^n v, so <lam, v> = ||J^n v||^2
    a = float((lam * v).sum())
    b = float((Jv * Jv).sum())
    assert abs(a - b) <= 1e-11 * abs(b), (a, b)
    prop.close()


@pytest.mark.parametrize("stepping", [2, 3])
def test_linearised_steps_need_rk4(torch_cuda, stepping):
    """The tangent and adjoint steps are RK4's; other integrators are refused with an error."""
    from gatt_dns_amd import runtime
    c, q = _case("box3d")
    c.time_stepping = stepping
    s = runtime.Solver(c)
    s.set_state(q)
    with pytest.raises(runtime.GattError, match="RK4"):
        s.tlm_step(0.0, c.dt)
    with pytest.raises(runtime.GattError, match="RK4"):
        s.adj_step(0.0, c.dt)


@pytest.mark.parametrize("kind", ["box3d_forced", "bl3d", "bl_sfd"])
def test_members_linearised_alone(torch_cuda, kind):
    """Two members of one context (different base states and perturbations; the forcing's sums
    per member; an evolving SFD mean per member) advance their tangent and adjoint steps as each
    does alone, bit for bit."""
    from gatt_dns_amd import runtime
    case, q = _sfd_case(kind)[:2] if kind in SFD_KINDS else _case(kind)
    rng = np.random.Generator(np.random.PCG64(17))
    q2 = {k: np.asfortranarray(v * (1.0 + 1e-3 * rng.standard_normal(v.shape))) for k, v in q.items()}
    bases = [q, q2]
    perts = [{k: np.asfortranarray(rng.standard_normal(q[k].shape)) for k in "UVWRE"} for _ in range(2)]
    for step in ("tlm", "adj"):
        pair = runtime.Solver(case, batch=2)
        for m in range(2):
            pair.set_state(bases[m], member=m)
            pair.set_perturbation(perts[m], member=m)
        getattr(pair, step + "_step")(0.0, case.dt)
        together = [pair.get_perturbation(member=m) for m in range(2)]
        pair.close()
        for m in range(2):
            one = runtime.Solver(case)
            one.set_state(bases[m])
            one.set_perturbation(perts[m])
            getattr(one, step + "_step")(0.0, case.dt)
            alone = one.get_perturbation()
            one.close()
            for k in "UVWRE":
                assert np.array_equal(together[m][k], alone[k]), (kind, step, m, k)


def test_mean_slots(torch_cuda):
    """The SFD mean's copy slot: refused without SFD; with SFD the mean reads as the state until
    it is set or first used, and a mean set through the slot is the one the step starts from."""
    import torch
    from gatt_dns_amd import runtime
    box, q = _case("box2d")
    s = runtime.Solver(box)
    t = torch.zeros((5, box.nz, box.ny, box.nx), dtype=torch.float64, device="cuda")
    with pytest.raises(runtime.GattError, match="no SFD mean"):
        s.copy_fields(runtime.SLOT_MEAN, False, t)
    s.close()
    case, q, mean = _sfd_case("bl_sfd")
    from gatt_dns_amd.adjoint import state_tensor
    s = runtime.Solver(case)
    s.set_state(q)
    got = torch.empty_like(state_tensor(case, q))
    s.copy_fields(runtime.SLOT_MEAN, False, got)
    assert torch.equal(got[[0, 1, 3, 4]], state_tensor(case, q)[[0, 1, 3, 4]])
    m = state_tensor(case, mean)
    s.copy_fields(runtime.SLOT_MEAN, True, m)
    back = s.get_mean()
    for i, k in enumerate("UVWRE"):
        if k == "W":
            continue
        assert np.array_equal(back[k], mean[k]), k
    s.close()
