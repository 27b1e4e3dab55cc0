"""The fused stage (GATT_FUSED_RHS=1: x/z flux derivatives computed from E and the first
derivatives as they load, the y flux derivatives fused with the assembly and the stage
update) against the separate visc / second-derivative / assemble passes (the default) and
the oracle.

The two GPU paths apply the same operators; they differ by the rounding of the y solve's
segmented scan (8-point segments instead of 16) and the order of the forcing sums, so
they agree to 1e-12 per field (each field on its own scale) after three steps.  One step
is also checked against the oracle at 1e-12.
"""
import os

import numpy as np
import pytest

from gatt_dns_amd.prep import cases

pytestmark = pytest.mark.gpu

CASES = {
    "box32": lambda: cases.turbulence_box(32),
    "cavity33": lambda: cases.lid_driven(33),
    "bl160x64": lambda: cases.boundary_layer(nx=160, ny=64, nbuf=12),
    "bl64x40x12": lambda: cases.boundary_layer(nx=64, ny=40, nbuf=8, nz=12),
    "channel3d": lambda: cases.poiseuille(nz=8),
}


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd import runtime
    runtime.load_library()
    return runtime


def rel(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale else float(np.max(np.abs(a)))


def moving(case, amp=0.01):
    st = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
    rng = np.random.default_rng(11)
    for k in ("U", "V") + (("W",) if case.is3d else ()):
        st[k] = np.asfortranarray(st[k] + amp * rng.standard_normal(st[k].shape))
    if case.flow_region is not None:
        wall = ~np.asarray(case.flow_region, bool).reshape(st["U"].shape, order="F")
        for k in ("U", "V", "W"):
            st[k][wall] = 0.0
    return st


def make(rt, case, unfused):
    old = os.environ.get("GATT_FUSED_RHS")
    os.environ["GATT_FUSED_RHS"] = "0" if unfused else "1"
    try:
        return rt.Solver(case)
    finally:
        if old is None:
            del os.environ["GATT_FUSED_RHS"]
        else:
            os.environ["GATT_FUSED_RHS"] = old


@pytest.mark.parametrize("name", list(CASES))
def test_fused_equals_unfused(rt, oracle_mod, name):
    case = CASES[name]()
    print("fused vs unfused stage:", name, flush=True)
    q0 = moving(case)
    fused, plain = make(rt, case, False), make(rt, case, True)
    try:
        # the fused kernel class exists only where the step uses it
        fused.time_kernel(10, 1)
        with pytest.raises(rt.GattError):
            plain.time_kernel(10, 1)
        fused.set_state(q0)
        plain.set_state(q0)
        fused.apply_bound_cond(0.0)
        plain.apply_bound_cond(0.0)
        # one step against the oracle
        o = oracle_mod.Oracle(case)
        qo = o.apply_bound_cond(q0, 0.0)
        o.init_mean(qo)
        want, _, _ = o.run(qo, 0.0, case.dt, 1)
        fused.step(0.0, case.dt)
        got = fused.get_state()
        for k in ("UVWRE" if case.is3d else "UVRE"):
            assert rel(got[k], want[k]) < 1e-12, (name, k, rel(got[k], want[k]))
        plain.step(0.0, case.dt)
        t = case.dt
        fused.advance(2, t, case.dt)
        plain.advance(2, t, case.dt)
        a, b = fused.get_state(), plain.get_state()
        for k in ("UVWRE" if case.is3d else "UVRE"):
            assert rel(a[k], b[k]) < 1e-12, (name, k, rel(a[k], b[k]))
    finally:
        fused.close()
        plain.close()
