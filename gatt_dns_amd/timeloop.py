"""The main loop around the time step: time control, log file, flow and mean-flow files.

``main.F90:170-531``: each iteration picks dt (fixed, ``timeControl`` 1, or the
largest dt under ``maxCFL`` that divides the save interval, ``timeControl`` 2),
advances the flow one step, writes a line to ``log.txt`` every ``logAll`` steps
and at every save, writes ``flow_%010d.mat`` (and, with an evolving SFD mean,
``meanflowSFD.mat``) at every save, and stops at ``tmax`` or when the flow
diverges (a NaN or an all-zero change; then ``flow_diverged.mat``).

The steps between two log or save points run as one captured-graph replay
(``gatt_advance``); the last step of each stretch runs through ``gatt_step``,
which returns the max |U|, |V|, |W| the log's CFL needs.  The divergence flag of
``gatt_advance`` accumulates over the stretch; when it is raised, the stretch is
run again one step at a time from a device checkpoint taken before it, so the
loop stops at the step that diverged and writes the state before that step, as
the reference does.
"""
import os

import numpy as np

from . import flowio
from .prep.cases import cfl_dt


def _cfl_dt_from_umax(umax, Ma, dxmin, nz):
    """CFLdt of main.F90:207-214 from max |U|, |V|, |W|."""
    c = (1 / Ma + umax[0]) * dxmin[0] + (1 / Ma + umax[1]) * dxmin[1]
    if nz > 1:
        c += (1 / Ma + umax[2]) * dxmin[2]
    return c


def _advance_time(t, dt, n):
    """t after n steps, accumulated one step at a time as main.F90:424 and the device clock do."""
    for _ in range(n):
        t += dt
    return t


class TimeLoop:
    """One run of the reference's main loop on a :class:`gatt_dns_amd.runtime.Solver`.

    ``time_control`` 1: fixed ``dt``; a save every ``qtimes`` steps, the run ends after
    step ``tmax`` (both step counts, ``qtimesInt`` / ``tmaxInt``).  ``time_control`` 2: dt
    from ``max_cfl`` capped at ``dt``, a save every ``qtimes`` time units, the run ends
    at time ``tmax``.  ``tracked``: rows of 0-based (i, j, k) grid points whose values
    go into every log line; ``tracked_norm`` > 0 logs R - 1 and E/tracked_norm - 1
    (``mesh.trackedNorm``, writeFortranParameters.m:27-31).  ``resume_mean``: start the SFD mean
    from ``meanflowSFD.mat`` in ``outdir`` (``resumeMeanFlow``); ``save_mean`` (default: SFD with
    Delta > 0) writes it at every save.
    """

    def __init__(self, solver, case, outdir, time_control=1, dt=None, qtimes=1, tmax=1, max_cfl=0.5, log_all=1,
                 tracked=(), tracked_norm=0.0, save_mean=None, resume_mean=False):
        self.s, self.case, self.outdir = solver, case, outdir
        self.time_control = int(time_control)
        self.dtmax = float(case.dt if dt is None else dt)
        self.qtimes, self.tmax = qtimes, tmax
        self.max_cfl = float(max_cfl)
        self.log_all = max(1, int(log_all))
        self.tracked = np.asarray(tracked, dtype=np.int64).reshape(-1, 3)
        self.tracked_norm = float(tracked_norm)
        self.save_mean = (case.sfd_type > 0 and case.sfd_delta > 0) if save_mean is None else save_mean
        self.resume_mean = bool(resume_mean)
        self.shape = (case.nx, case.ny, case.nz)
        self.log_path = os.path.join(outdir, "log.txt")
        if self.time_control not in (1, 2):
            raise ValueError("time_control must be 1 (fixed dt) or 2 (maximum CFL)")

    # -- outputs ---------------------------------------------------------------
    def _log(self, nsave_tmp, tstep, t, dt, cfl, max_change):
        vals = self.s.get_points(self.tracked) if len(self.tracked) else ()
        line = flowio.log_line(nsave_tmp, tstep, t, dt, cfl, max_change, vals, self.tracked_norm)
        if not os.path.exists(self.log_path):
            flowio.append_log(self.log_path, flowio.log_header(len(self.tracked)))
        flowio.append_log(self.log_path, line)

    def _save(self, nsave, t):
        fr = self.case.flow_region
        flowio.write_flow(os.path.join(self.outdir, flowio.flow_file_name(nsave)), self.s.get_state(), t,
                          self.shape, flow_region=fr)
        if self.save_mean and nsave >= 0:
            flowio.write_mean_flow(os.path.join(self.outdir, "meanflowSFD.mat"), self.s.get_mean(), self.shape,
                                   flow_region=fr)

    def _save_diverged(self, t):
        """flow_diverged.mat: the state before the diverged step, with the time after it
        (exportFlow runs before U = Unew, main.F90:423-441)."""
        fr = self.case.flow_region
        flowio.write_flow(os.path.join(self.outdir, flowio.flow_file_name(-1)), self.s.get_previous_state(), t,
                          self.shape, flow_region=fr)

    def _stretch(self, n, t, dt):
        """n steps: n - 1 replayed, the last through gatt_step.  Returns (steps run, umax,
        max_change, diverged).  A divergence inside the replayed part is found again one
        step at a time from a checkpoint taken before it, so the loop stops at the step
        that diverged, as main.F90:433-448 does."""
        if n > 1:
            self.s.checkpoint(0)
            _, dv = self.s.advance(n - 1, t, dt)
            if dv:
                self.s.checkpoint(1)
                tt = t
                for k in range(1, n + 1):
                    um, mc, dv = self.s.step(tt, dt)
                    tt += dt
                    if dv:
                        return k, um, mc, True
                return n, um, mc, False
        um, mc, dv = self.s.step(_advance_time(t, dt, n - 1), dt)
        return n, um, mc, dv

    # -- the loop --------------------------------------------------------------
    def run(self, t0=0.0, nsave0=0):
        """Run from time ``t0`` and save index ``nsave0`` (the flow file the run starts
        from); returns (t, tstep, nsave, diverged)."""
        os.makedirs(self.outdir, exist_ok=True)
        c = self.case
        if self.resume_mean:
            # resumeMeanFlow = 1 (main.F90:150-158, importMeanFlow.F90): the mean of the earlier run
            wall = None if c.flow_region is None else ~np.asarray(c.flow_region, bool).reshape(self.shape, order="F")
            self.s.set_mean(flowio.read_mean_flow(os.path.join(self.outdir, "meanflowSFD.mat"), self.shape, wall))
        t, nsave = float(t0), int(nsave0)
        # the boundary conditions on the imported flow before the loop (main.F90:110-117); a
        # held inflow (holdInlet) is captured here
        self.s.apply_bound_cond(t)
        if self.time_control == 1:
            tstep = nsave                                                # main.F90:170-172
            qt, tmax = int(self.qtimes), int(self.tmax)
            dt = self.dtmax
            while True:
                # steps to the next log point, save point or the end (main.F90:218-229)
                stop = min(tmax, (tstep // qt + 1) * qt, (tstep // self.log_all + 1) * self.log_all)
                n = max(1, stop - tstep)
                n, um, mc, diverged = self._stretch(n, t, dt)
                tstep += n
                t = _advance_time(t, dt, n)
                cfl = dt * _cfl_dt_from_umax(um, c.Ma, c.dxmin, c.nz)
                if diverged:                                             # main.F90:434-444
                    self._save_diverged(t)
                    return t, tstep, -1, True
                saving = tstep % qt == 0
                if saving:
                    nsave = tstep
                if tstep % self.log_all == 0 or saving:
                    self._log(nsave if saving else 0, tstep, t, dt, cfl, mc)
                if saving:
                    self._save(nsave, t)
                if tstep >= tmax:
                    return t, tstep, nsave, False
        tstep = 0                                                        # main.F90:173-174
        qt, tmax = float(self.qtimes), float(self.tmax)
        while True:
            # a save interval: dt from the CFL of its first state (main.F90:232-246)
            cdt = cfl_dt(self.s.get_state(), c.Ma, c.dxmin, c.nz)
            dt = min(self.max_cfl / cdt, self.dtmax)
            steps = int(np.floor(qt / dt)) + 1
            dt = qt / steps
            nsave += 1
            left = steps
            while left > 0:
                n = min(left, self.log_all - tstep % self.log_all)
                # the run ends with the first step that starts at t >= tmax - dt (main.F90:244-246)
                tt, last_in_run = t, False
                for j in range(1, n + 1):
                    if tt + dt >= tmax:
                        n, last_in_run = j, True
                        break
                    tt += dt
                n, um, mc, diverged = self._stretch(n, t, dt)
                tstep += n
                t = _advance_time(t, dt, n)
                left -= n
                cfl = dt * _cfl_dt_from_umax(um, c.Ma, c.dxmin, c.nz)
                if diverged:
                    self._save_diverged(t)
                    return t, tstep, -1, True
                saving = left == 0
                if tstep % self.log_all == 0 or saving:
                    self._log(nsave if saving else 0, tstep, t, dt, cfl, mc)
                if saving:
                    self._save(nsave, t)
                if last_in_run:
                    return t, tstep, nsave, False
