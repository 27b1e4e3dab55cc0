"""Ragged groups: members of different grids advanced together equal each member advanced alone."""
import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nsteps", [4, 5])
def test_group_equals_members_alone(mode, nsteps):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd.runtime import Group, Solver
    sweep = cases.cavity_sweep(3, nx_range=(33, 49), ny_range=(17, 33), re_range=(500.0, 1500.0))
    solo, grouped = [], []
    for c in sweep:
        for lst in (solo, grouped):
            s = Solver(c)
            s.set_state(c.state0)
            s.apply_bound_cond(0.0)
            lst.append(s)
    for c, s in zip(sweep, solo):
        s.advance(nsteps, 0.0, c.dt)
    g = Group(grouped, mode=mode)
    mc, dv = g.advance(nsteps, 0.0, [c.dt for c in sweep])
    assert not any(dv)
    for m, (a, b) in enumerate(zip(solo, grouped)):
        qa, qb = a.get_state(), b.get_state()
        for k in "UVRE":
            np.testing.assert_array_equal(qb[k], qa[k])
        assert mc[m].max() > 0
    # a second call continues from the advanced states
    g.advance(2, nsteps * sweep[0].dt, [c.dt for c in sweep])
    for c, s in zip(sweep, solo):
        s.advance(2, nsteps * c.dt, c.dt)
    for a, b in zip(solo, grouped):
        qa, qb = a.get_state(), b.get_state()
        for k in "UVRE":
            np.testing.assert_array_equal(qb[k], qa[k])
    g.close()


def _shapes(solver, ndir):
    return {(d, k): solver.get_line_shape(d, k) for d in range(ndir) for k in (0, 1)}


def _alone(case, shapes, nsteps):
    """One member advanced on its own with the launch shapes it had inside the group."""
    from gatt_dns_amd.runtime import Solver
    s = Solver(case)
    for (d, k), (v, G, M, w) in shapes.items():
        s.set_line_shape(d, k, v, G if v == 0 else M, w)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    mc, dv = s.advance(nsteps, 0.0, case.dt)
    return s, mc


@pytest.mark.parametrize("nsteps", [4, 5])
def test_ragged_launch_equals_members_alone(nsteps):
    """Mode 2 (every member in each launch, one graph of ragged launches): cavities of different
    sizes and Reynolds numbers whose x lines take the one-wavefront solve, each equal bit for bit
    to the member advanced alone with the same launch shapes; the members get their own shapes
    back when the group is destroyed."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd.runtime import Group, Solver
    members = cases.cavity_sweep(3, nx_range=(256, 520), ny_range=(40, 72), re_range=(700.0, 1300.0))
    grouped = []
    for c in members:
        s = Solver(c)
        s.set_state(c.state0)
        s.apply_bound_cond(0.0)
        grouped.append(s)
    own = [_shapes(s, 2) for s in grouped]
    g = Group(grouped, mode=2)
    shapes = [_shapes(s, 2) for s in grouped]
    # one wavefront per x line in every member, each with the fewest points per lane that cover its
    # lines (the ragged form of that kernel takes them per member)
    for m, c in enumerate(members):
        v, G, M, w = shapes[m][(0, 0)]
        assert v == 4 and 64 * M >= c.nx and (M == 4 or 32 * M < c.nx), (m, shapes[m][(0, 0)])
    assert shapes[0][(0, 0)][2] < shapes[2][(0, 0)][2]
    print("ragged group shapes", shapes, flush=True)
    mc, dv = g.advance(nsteps, 0.0, [c.dt for c in members])
    assert not any(dv)
    for m, c in enumerate(members):
        s1, mc1 = _alone(c, shapes[m], nsteps)
        qa, qb = s1.get_state(), grouped[m].get_state()
        for k in "UVRE":
            np.testing.assert_array_equal(qb[k], qa[k])
        np.testing.assert_array_equal(mc[m], mc1)
        s1.close()
    g.close()
    assert [_shapes(s, 2) for s in grouped] == own


@pytest.mark.parametrize("other", ["boundary_layer", "channel"])
def test_ragged_launch_refuses_different_steps(other):
    """A member whose step has other launches (the boundary layer's SFD pass) or other kernels
    (the periodic channel's x band is narrower than the cavity's, whose wall closures widen it)
    cannot share the ragged launches of a cavity: the advance reports it instead of running."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd.runtime import Group, Solver
    second = cases.boundary_layer(nx=64, ny=40, nbuf=8) if other == "boundary_layer" else \
        cases.poiseuille(nx=64, ny=48, spatial="SL4")
    members = [cases.lid_driven(33), second]
    solvers = []
    for c in members:
        s = Solver(c)
        s.set_state(c.state0)
        s.apply_bound_cond(0.0)
        solvers.append(s)
    g = Group(solvers, mode=2)
    with pytest.raises(RuntimeError, match="differ"):
        g.advance(2, 0.0, [c.dt for c in members])
    g.close()


@pytest.mark.parametrize("kind", ["boxes", "boundary_layers"])
def test_ragged_launch_3d_members(kind):
    """Mode 2 with 3-D members: turbulence boxes of 16^3, 24^3 and 32^3 (their lines take one
    lane-group shape of one point count per lane, each member its own lanes per line) and
    spanwise-periodic boundary layers of two lengths, each equal bit for bit to the member
    advanced alone with the same launch shapes."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd.runtime import Group, Solver
    if kind == "boxes":
        members = [cases.turbulence_box(n, seed=3 + n, forcing=False) for n in (16, 24, 32)]
    else:
        members = [cases.boundary_layer(nx=nx, ny=40, nbuf=8, nz=12, noise=0.0) for nx in (64, 96)]
    grouped = []
    for c in members:
        s = Solver(c)
        s.set_state(c.state0)
        s.apply_bound_cond(0.0)
        grouped.append(s)
    own = [_shapes(s, 3) for s in grouped]
    g = Group(grouped, mode=2)
    shapes = [_shapes(s, 3) for s in grouped]
    print("3-D ragged group", kind, shapes, flush=True)
    nsteps = 4
    mc, dv = g.advance(nsteps, 0.0, [c.dt for c in members])
    assert not any(dv)
    for m, c in enumerate(members):
        s1 = Solver(c)
        for (d, k), (v, G, M, w) in shapes[m].items():
            s1.set_line_shape(d, k, v, G if v == 0 else M, w)
        s1.set_state(c.state0)
        s1.apply_bound_cond(0.0)
        mc1, _ = s1.advance(nsteps, 0.0, c.dt)
        qa, qb = s1.get_state(), grouped[m].get_state()
        for k in "UVWRE":
            np.testing.assert_array_equal(qb[k], qa[k])
        np.testing.assert_array_equal(mc[m], mc1)
        s1.close()
    g.close()
    assert [_shapes(s, 3) for s in grouped] == own
