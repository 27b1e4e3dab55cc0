"""The main loop around the step (main.F90:170-531) and the tracked-point gather.

* ``gatt_get_points`` returns exactly the state values at the requested points
  (W = 0 in 2-D) and rejects points outside the grid.
* Fixed dt (timeControl 1): the loop's captured stretches plus the final
  ``gatt_step`` give the state of the same number of single steps bit for bit;
  flow files appear at the save steps and hold the state of that step; the log
  has the header and one line per ``logAll`` step and per save, with the save
  number only on save lines, the time accumulated step by step, the CFL of the
  state before the step, and the tracked values of the state after it.
* Maximum CFL (timeControl 2): dt divides the save interval, stays under
  maxCFL / CFLdt of the interval's first state, and the run ends with the first
  step that reaches tmax.
"""
import os

import numpy as np
import pytest

from gatt_dns_amd.prep import cases


@pytest.fixture(scope="module")
def rt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gatt_dns_amd import runtime
    runtime.load_library()
    return runtime


@pytest.mark.gpu
@pytest.mark.parametrize("mk", [lambda: cases.lid_driven(33), lambda: cases.turbulence_box(16)], ids=["2d", "3d"])
def test_get_points(rt, mk):
    case = mk()
    s = rt.Solver(case)
    try:
        s.set_state(case.state0)
        s.step(0.0, case.dt)
        st = s.get_state()
        rng = np.random.default_rng(3)
        pts = np.stack([rng.integers(0, n, 40) for n in (case.nx, case.ny, case.nz)], axis=1)
        got = s.get_points(pts)
        for f, k in enumerate("UVWRE"):
            want = st[k][pts[:, 0], pts[:, 1], pts[:, 2]]
            np.testing.assert_array_equal(got[:, f], want if case.is3d or k != "W" else 0.0)
        assert s.get_points(np.zeros((0, 3))).shape == (0, 5)
        with pytest.raises(rt.GattError):
            s.get_points([[case.nx, 0, 0]])
    finally:
        s.close()


def _moving(case, amp=0.02):
    """The initial state plus a smooth velocity perturbation, so that every step changes the flow."""
    st = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
    x = np.linspace(0, 1, case.nx)[:, None, None]
    y = np.linspace(0, 1, case.ny)[None, :, None]
    st["U"] = np.asfortranarray(st["U"] + amp * np.sin(np.pi * x) * np.sin(2 * np.pi * y))
    st["V"] = np.asfortranarray(st["V"] + amp * np.sin(2 * np.pi * x) * np.sin(np.pi * y))
    return st


def _log_rows(path):
    lines = open(path).read().split("\n")
    return lines[0], [l for l in lines[1:] if l]


@pytest.mark.gpu
def test_fixed_dt_loop(rt, tmp_path):
    from gatt_dns_amd import flowio
    from gatt_dns_amd.timeloop import TimeLoop
    case = cases.lid_driven(33)
    tracked = [[3, 4, 0], [16, 30, 0]]
    s = rt.Solver(case)
    ref = rt.Solver(case)
    try:
        q0 = _moving(case)
        s.set_state(q0)
        ref.set_state(q0)
        loop = TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=5, tmax=12, log_all=2, tracked=tracked,
                        tracked_norm=case.E0)
        t, tstep, nsave, diverged = loop.run()
        assert (tstep, nsave, diverged) == (12, 10, False)
        # the same 12 steps one at a time, after the loop's initial boundary application
        ref.apply_bound_cond(0.0)
        tr, rows, saved = 0.0, {}, {}
        for k in range(1, 13):
            um, mc, dv = ref.step(tr, case.dt)
            tr += case.dt
            rows[k] = (um, mc, ref.get_points(tracked))
            if k % 5 == 0:
                saved[k] = ref.get_state()
        assert t == tr
        got, want = s.get_state(), ref.get_state()
        for k in "UVRE":
            np.testing.assert_array_equal(got[k], want[k])
        for k in (5, 10):
            q, tf, _ = flowio.read_flow(os.path.join(tmp_path, flowio.flow_file_name(k)), loop.shape)
            assert tf == pytest.approx(k * case.dt, rel=1e-15)
            fr = case.flow_region.reshape(loop.shape, order="F")      # walls come back as 0 / 1 by design
            for f in "UVRE":
                np.testing.assert_array_equal(q[f][fr], saved[k][f][fr])
        assert not os.path.exists(os.path.join(tmp_path, flowio.flow_file_name(12)))
        header, lines = _log_rows(os.path.join(tmp_path, "log.txt"))
        assert header == flowio.log_header(2).rstrip("\n")
        steps = [int(l.split("\t")[1]) for l in lines]
        assert steps == [2, 4, 5, 6, 8, 10, 12]
        tt = 0.0
        times = {}
        for k in range(1, 13):
            tt += case.dt
            times[k] = tt
        for l, k in zip(lines, steps):
            um, mc, vals = rows[k]
            cfl = case.dt * ((1 / case.Ma + um[0]) * case.dxmin[0] + (1 / case.Ma + um[1]) * case.dxmin[1])
            want = flowio.log_line(k if k % 5 == 0 else 0, k, times[k], case.dt, cfl, mc, vals, case.E0)
            assert l + "\n" == want, (k, l, want)
    finally:
        s.close()
        ref.close()


@pytest.mark.gpu
def test_cfl_loop(rt, tmp_path):
    from gatt_dns_amd.prep.cases import cfl_dt
    from gatt_dns_amd.timeloop import TimeLoop
    case = cases.lid_driven(33)
    s = rt.Solver(case)
    try:
        q0 = _moving(case)
        # the first interval's dt comes from the state after the loop's initial boundary
        # application (main.F90:110-117, then 232-246): the lid moves at U0
        s.set_state(q0)
        s.apply_bound_cond(0.0)
        cdt = cfl_dt(s.get_state(), case.Ma, case.dxmin, case.nz)
        s.set_state(q0)
        qtimes = 3.3 * case.dt
        loop = TimeLoop(s, case, str(tmp_path), time_control=2, dt=10 * case.dt, qtimes=qtimes, tmax=2 * qtimes,
                        max_cfl=0.5, log_all=1000)
        t, tstep, nsave, diverged = loop.run()
        dt = min(0.5 / cdt, 10 * case.dt)
        steps = int(np.floor(qtimes / dt)) + 1
        assert qtimes / steps <= 0.5 / cdt
        assert not diverged and nsave >= 2 and t >= 2 * qtimes - 1e-12
        lines = open(os.path.join(tmp_path, "log.txt")).read().split("\n")[1:]
        lines = [l for l in lines if l]
        assert int(lines[0].split("\t")[0]) == 1 and int(lines[0].split("\t")[1]) == steps
        # the second interval takes its dt from its own first state; t may land an ulp short of
        # tmax after it, and then one more step ends the run
        assert int(lines[1].split("\t")[0]) == 2 and int(lines[1].split("\t")[1]) in (tstep, tstep - 1)
    finally:
        s.close()


@pytest.mark.gpu
def test_automean(rt, tmp_path):
    """runDNSautomean.m: the averaged flow file, the log2 lines, the continued runs and the stop
    after n_means passes; the averaged file is the in-order sum of the last files over their count."""
    from gatt_dns_amd import automean, flowio
    from gatt_dns_amd.timeloop import TimeLoop
    case = cases.lid_driven(33)
    shape = (case.nx, case.ny, case.nz)
    s = rt.Solver(case)
    try:
        s.set_state(_moving(case))
        TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=2, tmax=8, log_all=1).run()
        assert automean.flow_files(str(tmp_path))[-1] == flowio.flow_file_name(8)
        passes = automean.run_automean(s, case, str(tmp_path), files_for_mean=3, spacing_for_mean=2, n_means=2,
                                       tol=0.0, time_control=1, qtimes=2, log_all=1)
        assert passes == 2
        names = automean.flow_files(str(tmp_path))
        steps = [int(n[5:15]) for n in names]
        assert steps == [2, 4, 6, 8, 9, 10, 12, 13, 14, 16]
        log2 = open(os.path.join(tmp_path, "log2.txt")).read().split("\n")
        assert log2[:2] == ["Mean flow calculated from flow_0000000004.mat to flow_0000000008.mat",
                            "Mean flow calculated from flow_0000000009.mat to flow_0000000012.mat"]
        fr = case.flow_region.reshape(shape, order="F")
        parts = [flowio.read_flow(os.path.join(tmp_path, flowio.flow_file_name(k)), shape) for k in (4, 6, 8)]
        mean, t9, _ = flowio.read_flow(os.path.join(tmp_path, flowio.flow_file_name(9)), shape)
        assert t9 == parts[-1][1]
        for k in "UVRE":
            want = ((0.0 + parts[0][0][k]) + parts[1][0][k]) + parts[2][0][k]
            np.testing.assert_array_equal(mean[k][fr], (want / 3)[fr])
        # the last log line is the last step of the second continued run
        assert int(open(os.path.join(tmp_path, "log.txt")).read().strip().split("\n")[-1].split("\t")[1]) == 16
        assert automean.last_change(os.path.join(tmp_path, "log.txt")) > 0
    finally:
        s.close()


def _rel(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale else float(np.max(np.abs(a)))


@pytest.mark.gpu
def test_held_inflow_loop(rt, oracle_mod, tmp_path):
    """A boundary layer that starts past the leading edge holds its inflow plane (holdInlet,
    boundaryLayerIsothermal.m:53-67).  The loop's boundary application on the imported flow
    (main.F90:110-117) captures it from the initial flow, also when every stretch is one step;
    the result matches the oracle run with its own initial application, field by field."""
    from gatt_dns_amd.timeloop import TimeLoop
    case = cases.boundary_layer(nx=96, ny=48, nbuf=10, x_range=(30.0, 60.0))
    assert [d.kind for d in case.disturb] == ["holdInlet"]
    print("held-inflow boundary layer 96x48, 3 one-step stretches", flush=True)
    o = oracle_mod.Oracle(case)
    q0 = o.apply_bound_cond(case.state0, 0.0)
    o.init_mean(q0)
    want, _, _ = o.run(q0, 0.0, case.dt, 3)
    s = rt.Solver(case)
    try:
        s.set_state(case.state0)
        t, tstep, _, diverged = TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=100, tmax=3,
                                         log_all=1).run()
        assert tstep == 3 and not diverged
        got = s.get_state()
        # each field on its own scale: V (about 1e-3) carries the round-off of the O(1)
        # pressure terms, about 2e-14 absolute after three steps
        for k in "UVRE":
            assert _rel(got[k], want[k]) < 1e-10, (k, _rel(got[k], want[k]))
    finally:
        s.close()


@pytest.mark.gpu
def test_divergence_flag_per_call(rt):
    """The divergence flag describes the steps of one call: a step whose change is NaN raises
    it (main.F90:434), and later steps of a finite flow on the same context do not report it
    again."""
    case = cases.periodic_rest(16, 16)
    s = rt.Solver(case)
    try:
        bad = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
        bad["U"][3, 5, 0] = np.nan
        s.set_state(bad)
        _, mc, dv = s.step(0.0, case.dt)
        assert dv and np.isnan(mc[0])
        moving = {k: np.array(v, order="F", copy=True) for k, v in case.state0.items()}
        x = np.arange(case.nx)[:, None, None] * (2 * np.pi / case.nx)
        moving["U"] = np.asfortranarray(moving["U"] + 0.05 * np.sin(x) * np.ones(moving["U"].shape))
        s.set_state(moving)
        _, _, dv = s.step(0.0, case.dt)
        assert not dv
        _, dv = s.advance(4, 0.0, case.dt)
        assert not dv
    finally:
        s.close()


@pytest.mark.gpu
def test_checkpoint_and_previous_state(rt):
    """gatt_checkpoint restores the state (and the SFD mean) exactly, so a stretch repeated one
    step at a time from it equals the replayed stretch bit for bit; gatt_get_previous_state is
    the state before the last step."""
    case = cases.boundary_layer(nx=96, ny=48, nbuf=10)
    s = rt.Solver(case)
    try:
        s.set_state(case.state0)
        s.apply_bound_cond(0.0)
        before = s.get_state()
        s.step(0.0, case.dt)
        prev = s.get_previous_state()
        for k in "UVWRE":
            np.testing.assert_array_equal(prev[k], before[k])
        t1 = case.dt
        s.checkpoint(0)
        s.advance(3, t1, case.dt)
        replayed, mean_r = s.get_state(), s.get_mean()
        s.checkpoint(1)
        tt = t1
        for _ in range(3):
            s.step(tt, case.dt)
            tt += case.dt
        got, mean_g = s.get_state(), s.get_mean()
        for k in "UVRE":
            np.testing.assert_array_equal(got[k], replayed[k])
            np.testing.assert_array_equal(mean_g[k], mean_r[k])
    finally:
        s.close()
