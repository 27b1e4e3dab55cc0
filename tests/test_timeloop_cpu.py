"""Cadence and time control of the main loop (main.F90:170-531) with a stand-in solver.

The stand-in records every call and adds one to U per step, so the checks are
on the loop's own logic: which steps are replayed together, the times handed
to each call (accumulated one step at a time), the log and save points, the
save numbers, the CFL-controlled dt and the stop at tmax, the divergence stop
(flow_diverged.mat) and the automean file sequence.
"""
import os

import numpy as np
import pytest

from gatt_dns_amd import flowio


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(flowio.LIB_PATH):
        from gatt_dns_amd import build
        if build.build_flowio(verbose=False) is None:
            pytest.skip("no HDF5 headers")
    return flowio.load_library()


class Case:
    nx, ny, nz = 4, 3, 1
    Ma = 0.5
    dxmin = (10.0, 20.0, 0.0)
    dt = 0.01
    sfd_type = 0
    sfd_delta = 0.0
    flow_region = None


class FakeSolver:
    def __init__(self, case, diverge_at=None):
        self.shape = (case.nx, case.ny, case.nz)
        self.q = {k: np.full(self.shape, 1.0 if k in "RE" else 0.0, order="F") for k in "UVWRE"}
        self.calls = []
        self.bc = []
        self.steps = 0
        self.diverge_at = diverge_at

    def _one(self):
        self.prev = {k: v.copy() for k, v in self.q.items()}
        self.steps += 1
        self.q["U"] = self.q["U"] + 1.0
        return self.diverge_at is not None and self.steps >= self.diverge_at

    def apply_bound_cond(self, t):
        self.bc.append(t)

    def checkpoint(self, op):
        if op == 0:
            self.saved = ({k: v.copy() for k, v in self.q.items()}, self.steps)
        else:
            q, self.steps = self.saved
            self.q = {k: v.copy() for k, v in q.items()}

    def get_previous_state(self):
        return {k: v.copy() for k, v in self.prev.items()}

    def advance(self, n, t0, dt):
        self.calls.append(("advance", n, t0, dt))
        dv = False
        for _ in range(n):
            dv = self._one() or dv
        return np.full(5, 1e-3), dv

    def step(self, t, dt):
        self.calls.append(("step", 1, t, dt))
        um = np.array([np.max(np.abs(self.q["U"])), 0.0, 0.0])
        dv = self._one()
        return um, np.array([1e-3, 2e-3, 0.0, 4e-3, 5e-3]), dv

    def get_state(self):
        return {k: v.copy() for k, v in self.q.items()}

    def set_state(self, q):
        self.q = {k: np.array(q[k], order="F", copy=True) for k in "UVWRE"}

    def set_mean(self, q):
        self.mean = {k: np.array(q[k], copy=True) for k in "UVWRE"}

    def get_mean(self):
        return self.mean

    def get_points(self, pts):
        return np.array([[self.q[k][tuple(p)] for k in "UVWRE"] for p in pts])


def _log(path):
    return [l.split("\t") for l in open(path).read().split("\n")[1:] if l]


def test_fixed_dt_cadence(lib, tmp_path):
    from gatt_dns_amd.timeloop import TimeLoop
    case = Case()
    s = FakeSolver(case)
    loop = TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=5, tmax=12, log_all=3, tracked=[[1, 1, 0]])
    t, tstep, nsave, dv = loop.run()
    assert (tstep, nsave, dv) == (12, 10, False)
    # stretches end at the log points 3, 6, 9, 12 and the saves 5, 10
    ends = np.cumsum([n for kind, n, _, _ in s.calls])
    assert [int(e) for e, (kind, *_r) in zip(ends, s.calls) if kind == "step"] == [3, 5, 6, 9, 10, 12]
    tt = 0.0
    times = [tt]
    for _ in range(12):
        tt += case.dt
        times.append(tt)
    start = 0
    for kind, n, t0, dt in s.calls:
        assert t0 == times[start] and dt == case.dt
        start += n
    assert t == times[12]
    rows = _log(os.path.join(tmp_path, "log.txt"))
    assert [int(r[1]) for r in rows] == [3, 5, 6, 9, 10, 12]
    assert [int(r[0]) for r in rows] == [0, 5, 0, 0, 10, 0]
    # the tracked U after step k is k; the CFL column is dt * CFLdt of max|U| before the last step
    assert [float(r[10]) for r in rows] == [3, 5, 6, 9, 10, 12]
    cfl = case.dt * ((1 / case.Ma + 2.0) * case.dxmin[0] + (1 / case.Ma) * case.dxmin[1])
    assert rows[0][4] == flowio.log_line(0, 3, 0.0, 1.0, cfl, [0] * 5).split("\t")[4]
    assert sorted(os.listdir(tmp_path)) == ["flow_0000000005.mat", "flow_0000000010.mat", "log.txt"]
    q, tf, _ = flowio.read_flow(os.path.join(tmp_path, "flow_0000000010.mat"), loop.shape)
    assert tf == times[10] and np.all(q["U"] == 10.0)


def test_resume_and_divergence(lib, tmp_path):
    from gatt_dns_amd.timeloop import TimeLoop
    case = Case()
    s = FakeSolver(case, diverge_at=4)
    loop = TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=5, tmax=30, log_all=100)
    t, tstep, nsave, dv = loop.run(t0=1.0, nsave0=10)
    # resumed at step 10: the first stretch (to the save at 15) carries the divergence of its
    # 4th step; it is run again one step at a time from the checkpoint and stops at step 14
    assert dv and nsave == -1 and tstep == 14
    kinds = [c[0] for c in s.calls]
    assert kinds == ["advance", "step", "step", "step", "step"]
    # the boundary conditions on the imported flow, once, at the start time (main.F90:110-117)
    assert s.bc == [1.0]
    tt = 1.0
    for _ in range(4):
        tt += case.dt
    assert t == tt
    # flow_diverged.mat holds the state before the diverged step with the time after it
    q, tf, _ = flowio.read_flow(os.path.join(tmp_path, "flow_diverged.mat"), loop.shape)
    assert tf == tt and np.all(q["U"] == 3.0)
    assert not os.path.exists(os.path.join(tmp_path, "flow_0000000015.mat"))


def test_cfl_control(lib, tmp_path):
    from gatt_dns_amd.timeloop import TimeLoop
    case = Case()
    s = FakeSolver(case)
    qtimes, tmax = 0.05, 0.12
    loop = TimeLoop(s, case, str(tmp_path), time_control=2, dt=0.02, qtimes=qtimes, tmax=tmax, max_cfl=0.5,
                    log_all=1000)
    t, tstep, nsave, dv = loop.run()
    # each interval: dt = min(maxCFL/CFLdt, dtmax) then qtimes/(floor(qtimes/dt)+1)
    dts, n_int, u = [], [], 0.0
    tt = 0.0
    while True:
        cdt = (1 / case.Ma + u) * case.dxmin[0] + (1 / case.Ma) * case.dxmin[1]
        dt = min(0.5 / cdt, 0.02)
        steps = int(np.floor(qtimes / dt)) + 1
        dt = qtimes / steps
        done = False
        for j in range(steps):
            if tt + dt >= tmax:
                done = True
            tt += dt
            u += 1.0
            if done:
                n_int.append(j + 1)
                break
        dts.append(dt)
        if done:
            break
        n_int.append(steps)
    assert tstep == sum(n_int) and t == tt and nsave == len(n_int)
    assert [c[3] for c in s.calls] == [d for d, n in zip(dts, n_int) for _ in range(2 if n > 1 else 1)]
    rows = _log(os.path.join(tmp_path, "log.txt"))
    assert [int(r[0]) for r in rows] == list(range(1, len(n_int)))     # the last interval ends at tmax, unsaved


def test_automean_files(lib, tmp_path):
    from gatt_dns_amd import automean
    from gatt_dns_amd.timeloop import TimeLoop
    case = Case()
    s = FakeSolver(case)
    TimeLoop(s, case, str(tmp_path), time_control=1, qtimes=2, tmax=8, log_all=1).run()
    assert automean.last_change(os.path.join(tmp_path, "log.txt")) == 5e-3
    n = automean.run_automean(s, case, str(tmp_path), files_for_mean=3, spacing_for_mean=2, n_means=1, tol=0.0,
                              qtimes=2)
    assert n == 1
    q, t, _ = flowio.read_flow(os.path.join(tmp_path, "flow_0000000009.mat"), (4, 3, 1))
    assert np.all(q["U"] == 6.0)                       # mean of the U = 4, 6, 8 files
    assert [f for f in automean.flow_files(str(tmp_path))][-3:] == ["flow_0000000009.mat", "flow_0000000010.mat",
                                                                   "flow_0000000012.mat"]
    q, _, _ = flowio.read_flow(os.path.join(tmp_path, "flow_0000000012.mat"), (4, 3, 1))
    assert np.all(q["U"] == 9.0)                       # restarted from the mean: 6 + 3 steps
    assert automean.run_automean(s, case, str(tmp_path), 3, 2, n_means=5, tol=1.0) == 0   # change below tol


def test_mean_flow_resume_and_save(lib, tmp_path):
    """resumeMeanFlow: the mean starts from meanflowSFD.mat (walls as importMeanFlow sets them);
    with SFD Delta > 0 the mean is written at every save."""
    from gatt_dns_amd.timeloop import TimeLoop
    case = Case()
    case.sfd_type, case.sfd_delta = 1, 0.5
    fr = np.ones((4, 3, 1), bool)
    fr[0] = False
    case.flow_region = fr
    mean = {k: np.full((4, 3, 1), v) for k, v in zip("UVWRE", (0.5, 0.25, 0.0, 1.5, 2.5))}
    flowio.write_mean_flow(os.path.join(tmp_path, "meanflowSFD.mat"), mean, (4, 3, 1), flow_region=fr)
    s = FakeSolver(case)
    TimeLoop(s, case, str(tmp_path), qtimes=2, tmax=2, resume_mean=True).run()
    assert np.all(s.mean["U"][fr] == 0.5) and np.all(s.mean["U"][~fr] == 0.0) and np.all(s.mean["E"][~fr] == 1.0)
    # a later run without resume writes the solver's current mean at its save
    s.set_mean({k: np.full((4, 3, 1), 7.0) for k in "UVWRE"})
    TimeLoop(s, case, str(tmp_path), qtimes=2, tmax=4).run(t0=0.02, nsave0=2)
    got = flowio.read_mean_flow(os.path.join(tmp_path, "meanflowSFD.mat"), (4, 3, 1))
    assert np.all(got["U"][fr] == 7.0) and np.all(got["U"][~fr] == 0.0)
