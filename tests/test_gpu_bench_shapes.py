"""Parity of the launch shapes and sizes the bench runs.

* Line solves at N = 256 along x, y and z with the shapes the 256^3 step uses (x: the
  32-line LDS tile with 16 points per lane; y and z: 16 segments of 16 points for the
  derivatives and 8 of 32 for the filter), against the oracle, ≤ 1e-13.
* The one-wavefront x solve with 8 points per lane on lines of exactly 64 x 8 = 512
  points (the x lines of the 512 x 192 x 64 boundary layer), against the oracle.
* One main-loop step of the 256^3 forced box and 1000 steps of the 1024 x 256 boundary
  layer against golden fixtures the oracle wrote (tests/golden/make_fixtures.py):
  sampled points measured against each field's own largest magnitude, ≤ 1e-12 for the
  step and ≤ 1e-10 after 1000 steps (V against its own scale, 1e-3 of U), and the
  max changes of the last step (1e-8 relative for the step: the change is a difference of
  O(1) values 1e-5 apart).
* Periodic solves and five steps on a non-uniform periodic mesh, where dX/deta differs
  between the first and the last node and the periodic solve's index shift of B and
  row-2 coupling through A(1) (derivs.F90:786-802) change the result.
"""
import os

import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd import runtime
    runtime.load_library()
    return runtime


def rel_err(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale else float(np.max(np.abs(a)))


def solve_and_compare(rt, s, o, case, d, kind, F, tol=1e-13):
    import torch
    shape = (case.nx, case.ny, case.nz)
    want = o.solve(d, kind, F)
    tin = torch.from_numpy(F.ravel(order="F").copy()).cuda()
    tout = torch.zeros_like(tin)
    s.line_solve(d, kind, [tin], [tout])
    got = tout.cpu().numpy().reshape(shape, order="F")
    err = rel_err(got, want)
    assert err < tol, (case.name, d, kind, s.get_line_shape(d, kind), err)
    return err


def test_line_solves_256_bench_shapes(rt, oracle_mod):
    case = cases.turbulence_box(256)
    o = oracle_mod.Oracle(case)
    s = rt.Solver(case)
    try:
        # the shapes the bench's 256^3 step runs (the defaults; the autotuner keeps them)
        expect = {(0, 0): (3, 16), (0, 1): (3, 16), (1, 0): (2, 16), (1, 1): (2, 32), (2, 0): (2, 16), (2, 1): (2, 32)}
        rng = np.random.default_rng(256)
        for (d, kind), (v, m) in expect.items():
            got_v, _, got_m, got_w = s.get_line_shape(d, kind)
            assert (got_v, got_m) == (v, m), (d, kind, got_v, got_m)
            if d == 0:
                assert got_w == 2, got_w   # the 16-line tile
            F = np.asfortranarray(rng.standard_normal((256, 256, 256)))
            print("256^3 line solve dir", d, "kind", kind, flush=True)
            solve_and_compare(rt, s, o, case, d, kind, F)
            # the 32-line tile along x as well
            if d == 0:
                s.set_line_shape(0, kind, 3, 16, 1)
                solve_and_compare(rt, s, o, case, d, kind, F)
                s.set_line_shape(0, kind, 3, 16, 2)
    finally:
        s.close()


def test_x_lines_512_one_wavefront(rt, oracle_mod):
    case = cases.boundary_layer(nx=512, ny=48, nbuf=20)
    o = oracle_mod.Oracle(case)
    s = rt.Solver(case)
    try:
        rng = np.random.default_rng(512)
        for kind in (0, 1):
            s.set_line_shape(0, kind, 4, 8, 1)
            assert s.get_line_shape(0, kind)[2] == 8
            F = np.asfortranarray(rng.standard_normal((512, 48, 1)))
            solve_and_compare(rt, s, o, case, 0, kind, F)
    finally:
        s.close()


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip("golden fixture %s not generated" % name)
    return np.load(path)


def _check_fixture(fx, case, got, mc, tol, tol_mc):
    check = np.array([np.sum(case.state0[k]) for k in "UVWRE"])
    np.testing.assert_allclose(check, fx["check"], rtol=1e-14, atol=1e-300)
    idx = fx["idx"]
    for f, k in enumerate("UVWRE"):
        if not case.is3d and k == "W":
            continue
        err = np.max(np.abs(np.asarray(got[k]).ravel(order="F")[idx] - fx["vals"][f])) / fx["qmax"][f]
        print(case.name, k, "sampled error / own max: %.3g" % err, flush=True)
        assert err <= tol, (k, err)
    for f in range(5):
        if fx["mc"][f]:
            assert abs(mc[f] - fx["mc"][f]) <= tol_mc * fx["mc"][f], (f, mc[f], fx["mc"][f])


def test_box256_step_against_fixture(rt):
    fx = _fixture("box256_step1.npz")
    case = cases.turbulence_box(256)
    s = rt.Solver(case)
    try:
        s.set_state(case.state0)
        _, mc, dv = s.step(0.0, case.dt)
        assert not dv
        # the max change is a difference of O(1) values 1e-5 apart: its relative error is the
        # fields' round-off over the change
        _check_fixture(fx, case, s.get_state(), mc, 1e-12, 1e-8)
    finally:
        s.close()


def test_boundary_layer_1000_steps_against_fixture(rt):
    fx = _fixture("bl1024x256_steps1000.npz")
    case = cases.boundary_layer()
    s = rt.Solver(case)
    try:
        s.set_state(case.state0)
        s.apply_bound_cond(0.0)
        print("boundary layer 1024x256, 1000 steps", flush=True)
        mc, dv = s.advance(int(fx["steps"]), 0.0, case.dt)
        assert not dv
        _check_fixture(fx, case, s.get_state(), mc, 1e-10, 1e-6)
    finally:
        s.close()


def test_nonuniform_periodic(rt, oracle_mod):
    case = cases.periodic_rest(24, 20, 16, stretch=0.3)
    o = oracle_mod.Oracle(case)
    s = rt.Solver(case)
    try:
        rng = np.random.default_rng(7)
        for d in range(3):
            for kind in (0, 1):
                F = np.asfortranarray(rng.standard_normal((24, 20, 16)))
                solve_and_compare(rt, s, o, case, d, kind, F)
        # a moving state, five steps
        q = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
        for k in "UVW":
            q[k] = np.asfortranarray(q[k] + 0.05 * rng.standard_normal(q[k].shape))
        want, _, dv_w = o.run(q, 0.0, case.dt, 5)
        s.set_state(q)
        mc, dv = s.advance(5, 0.0, case.dt)
        got = s.get_state()
        assert dv == dv_w
        for k in "UVWRE":
            assert rel_err(got[k], want[k]) <= 1e-11, (k, rel_err(got[k], want[k]))
    finally:
        s.close()
