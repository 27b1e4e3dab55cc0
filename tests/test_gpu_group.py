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
