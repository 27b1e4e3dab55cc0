"""Mean-flow iteration of ``runDNSautomean.m``.

Each pass (``runDNSautomean.m:33-108``):

1. reads the last line of ``log.txt`` and stops when the largest of its five
   max-change columns is below ``tol`` or ``n_means`` passes are done;
2. averages U, V, W, R, E over the last ``files_for_mean`` flow files (names
   ``flow_%010d.mat`` in name order), summing them one after the other from zero
   and dividing by the count, and writes the average, with the time of the last
   file, as ``flow_<last + 1>.mat``;
3. renames ``meanflowSFD.mat`` to ``meanflowSFD_<last>.mat`` (so the SFD mean
   starts again from the averaged flow);
4. continues the run from the averaged flow for ``spacing_for_mean`` save
   intervals (tmax = t + spacing * qtimes with CFL control, last + spacing *
   qtimes steps with fixed dt), and appends "Mean flow calculated from A to B" to
   ``log2.txt``.

The runs go through :class:`gatt_dns_amd.timeloop.TimeLoop` on a resident
solver instead of new processes.
"""
import os
import re

import numpy as np

from . import flowio
from .timeloop import TimeLoop

_FLOW = re.compile(r"^flow_\d{10}\.mat$")


def last_change(log_path):
    """max of the five change columns (6-10) of the last line of log.txt."""
    with open(log_path) as f:
        lines = [l for l in f.read().split("\n") if l.strip()]
    cols = lines[-1].split()
    return max(float(v) for v in cols[5:10])


def flow_files(outdir):
    return sorted(n for n in os.listdir(outdir) if _FLOW.match(n))


def average_flows(outdir, names, shape):
    """Sum of the files' U..E from zero, in order, divided by their count (runDNSautomean.m:59-77);
    NaN wall points stay NaN.  Returns (mean, t of the last file)."""
    acc = {k: 0.0 for k in "UVWRE"}
    t = None
    for n in names:
        q, t, wall = flowio.read_flow(os.path.join(outdir, n), shape)
        for k in "UVWRE":
            v = q[k].copy()
            v[wall] = np.nan
            acc[k] = acc[k] + v
    return {k: acc[k] / len(names) for k in "UVWRE"}, t


def run_automean(solver, case, outdir, files_for_mean, spacing_for_mean, n_means, tol, time_control=1, qtimes=1,
                 dt=None, max_cfl=0.5, log_all=1, tracked=(), tracked_norm=0.0, log=None):
    """The runDNSautomean.m loop on ``outdir``, which holds a finished run (flow files and log.txt).
    Returns the number of passes made."""
    shape = (case.nx, case.ny, case.nz)
    n_mean = 1
    while True:
        change = last_change(os.path.join(outdir, "log.txt"))
        if n_mean > n_means or change < tol:
            return n_mean - 1
        if log:
            log("Computing means loop: %d\n Current change: %g" % (n_mean, change))
        n_mean += 1
        names = flow_files(outdir)[-files_for_mean:]
        last = int(names[-1][5:15])
        mean, t = average_flows(outdir, names, shape)
        fr = None if case.flow_region is None else case.flow_region.reshape(shape, order="F")
        start = flowio.flow_file_name(last + 1)
        flowio.write_flow(os.path.join(outdir, start), mean, t, shape, flow_region=fr)
        mf = os.path.join(outdir, "meanflowSFD.mat")
        if os.path.exists(mf):
            os.replace(mf, os.path.join(outdir, "meanflowSFD_%d.mat" % last))
        with open(os.path.join(outdir, "log2.txt"), "a") as f:
            f.write("Mean flow calculated from %s to %s\n" % (names[0], names[-1]))
        # runDNS resumes from the newest flow file; the SFD mean starts from it (resumeMeanFlow = 0)
        q, t0, _ = flowio.read_flow(os.path.join(outdir, start), shape)
        solver.set_state(q)
        if case.sfd_type > 0:
            solver.set_mean(q)
        if time_control == 2:
            tmax = t + spacing_for_mean * qtimes
        else:
            tmax = last + spacing_for_mean * qtimes
        TimeLoop(solver, case, outdir, time_control=time_control, dt=dt, qtimes=qtimes, tmax=tmax, max_cfl=max_cfl,
                 log_all=log_all, tracked=tracked, tracked_norm=tracked_norm).run(t0=t0, nsave0=last + 1)
