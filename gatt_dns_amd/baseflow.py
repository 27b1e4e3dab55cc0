"""Steady base flows for the linearised analyses.

The reference obtains a steady base by running the DNS with selective frequency damping
(``main.F90:381-402``) or with the mean-flow iteration of ``runDNSautomean.m:33-108``
until the largest change of a step falls below a tolerance.  ``sfd_base`` does the
first on the GPU for the 2-D isothermal boundary layer (SFD type 1: the damping acts
everywhere) and stops after ``steps`` steps or when the largest change of a step is below
``tol``; ``extrude`` replicates a 2-D state along a spanwise-periodic z, which is the base
of the spanwise-periodic linearised problems (the 2-D base is a steady solution of the
3-D equations with W = 0 and no z dependence).
"""
import numpy as np

from .prep import cases
from .runtime import Solver


def sfd_base(nx=512, ny=192, steps=20000, chunk=5000, chi=0.1, delta=5.0, tol=0.0, log=None, **case_kw):
    """Run the 2-D boundary layer (cases.boundary_layer with no seeded noise) with SFD type 1.

    Returns ``(case, state, info)``: the 2-D case, its state after the run (field dict, nx x ny x 1
    arrays in column-major order) and ``{"steps", "t", "max_change"}`` (the largest change of each
    field over the last step)."""
    case = cases.boundary_layer(nx=nx, ny=ny, noise=0.0, sfd_type=1, sfd_chi=chi, sfd_delta=delta, **case_kw)
    s = Solver(case)
    try:
        s.set_state(case.state0)
        s.apply_bound_cond(0.0)
        done, t, mc = 0, 0.0, None
        while done < steps:
            n = min(chunk, steps - done)
            mc, dv = s.advance(n, t, case.dt)
            done += n
            t += n * case.dt
            if dv:
                raise RuntimeError("the SFD run diverged after %d steps" % done)
            if log:
                log("sfd base: %d steps, t = %.3f, max change %.3e" % (done, t, float(np.max(mc))))
            if float(np.max(mc)) < tol:
                break
        q = s.get_state()
    finally:
        s.close()
    return case, q, {"steps": done, "t": t, "max_change": [float(v) for v in mc]}


def extrude(q2d, nz):
    """A 2-D state (nx x ny x 1 fields) repeated along z (nz planes), W = 0."""
    out = {}
    for k in "UVWRE":
        a = np.asarray(q2d[k])
        a = a.reshape(a.shape[0], a.shape[1], -1)[:, :, :1]
        out[k] = np.asfortranarray(np.repeat(a, nz, axis=2))
    out["W"][:] = 0.0
    return out
