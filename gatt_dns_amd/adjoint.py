"""Adjoint of the time advance: a torch.autograd harness over the C ABI.

``advance(prop, q0, nsteps)`` returns q(T) = Phi^nsteps(q0) as a torch tensor
(shape (5, nz, ny, nx), fields U,V,W,R,E, each x fastest, float64 on the GPU).
Its backward pass runs the library's hand-written adjoint step
(``gatt_adj_step``) with RK-step checkpointing: the forward pass keeps every
``checkpoint_every``-th state on the GPU; the backward pass restores each
checkpoint, recomputes the states of its segment and applies the adjoint
steps in reverse.  The tangent (``gatt_tlm_step``) is exposed for
dot-product and finite-difference checks.

Where the SFD mean evolves (SFD on with Delta > 0, main.F90:382-388) the mean is
part of the dynamical state: the tensors then hold ten fields, the state's
U,V,W,R,E followed by the mean's, and every map (the step, its tangent, its
adjoint) acts on the pair.

Torch provides device memory and the autograd graph only; every arithmetic
operation of the step and of its adjoint runs in the library's HIP kernels.
"""
import numpy as np
import torch

from .runtime import SLOT_MEAN, SLOT_MEAN_PERTURBATION, SLOT_PERTURBATION, SLOT_STATE, Solver


def evolving_mean(case):
    """True where the SFD mean is part of the state (SFD on, Delta > 0)."""
    return case.sfd_type > 0 and case.sfd_delta > 0


class Propagator:
    """One-member context of a case with fixed dt; moves states between torch and the library.

    A state tensor has shape (5, nz, ny, nx), or (10, nz, ny, nx) where the SFD mean evolves
    (the mean's five fields after the state's)."""

    def __init__(self, case, dt=None, t0=0.0):
        self.case = case
        self.solver = Solver(case)
        self.dt = float(case.dt if dt is None else dt)
        self.t0 = float(t0)
        self.with_mean = evolving_mean(case)
        self.shape = (10 if self.with_mean else 5, case.nz, case.ny, case.nx)

    def empty(self):
        return torch.empty(self.shape, dtype=torch.float64, device="cuda")

    def _sync_torch(self):
        torch.cuda.current_stream().synchronize()

    def _copy(self, slots, to_ctx, t):
        self._sync_torch()
        self.solver.copy_fields(slots[0], to_ctx, t[:5])
        if self.with_mean:
            self.solver.copy_fields(slots[1], to_ctx, t[5:])

    def load(self, q):
        self._copy((SLOT_STATE, SLOT_MEAN), True, q.contiguous())

    def save(self):
        out = self.empty()
        self._copy((SLOT_STATE, SLOT_MEAN), False, out)
        return out

    def load_perturbation(self, v):
        self._copy((SLOT_PERTURBATION, SLOT_MEAN_PERTURBATION), True, v.contiguous())

    def save_perturbation(self):
        out = self.empty()
        self._copy((SLOT_PERTURBATION, SLOT_MEAN_PERTURBATION), False, out)
        return out

    def state_tensor(self, q, mean=None):
        """Host fields -> this propagator's state tensor; where the mean evolves, `mean` (default:
        the state itself, the mean the solver starts from) follows the state's fields."""
        t = state_tensor(self.case, q)
        if not self.with_mean:
            return t
        return torch.cat([t, state_tensor(self.case, q if mean is None else mean)])

    def steps(self, n, t=None):
        if n > 0:
            self.solver.advance(n, self.t0 if t is None else t, self.dt)

    def tangent(self, q, v, n=1):
        """(Phi^n(q), J v) with J the Jacobian of n steps at q."""
        self.load(q)
        self.load_perturbation(v)
        for s in range(n):
            self.solver.tlm_step(self.t0 + s * self.dt, self.dt)
        return self.save(), self.save_perturbation()

    def adjoint(self, q, w):
        """J(q)^T w for one step at q (q and w with the mean's fields where the mean evolves)."""
        self.load(q)
        self.load_perturbation(w)
        self.solver.adj_step(self.t0, self.dt)
        return self.save_perturbation()

    def close(self):
        self.solver.close()


class _Advance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, prop, nsteps, every):
        prop.load(q0.detach())
        checkpoints = []
        for s in range(0, nsteps, every):
            checkpoints.append(prop.save())
            prop.steps(min(every, nsteps - s), prop.t0 + s * prop.dt)
        out = prop.save()
        ctx.prop, ctx.nsteps, ctx.every, ctx.checkpoints = prop, nsteps, every, checkpoints
        return out

    @staticmethod
    def backward(ctx, gout):
        prop, nsteps, every = ctx.prop, ctx.nsteps, ctx.every
        lam = gout.detach().contiguous().clone()
        for seg in reversed(range(len(ctx.checkpoints))):
            start = seg * every
            length = min(every, nsteps - start)
            # recompute the segment's states from its checkpoint
            prop.load(ctx.checkpoints[seg])
            states = [ctx.checkpoints[seg]]
            for i in range(1, length):
                prop.steps(1, prop.t0 + (start + i - 1) * prop.dt)
                states.append(prop.save())
            for i in reversed(range(length)):
                prop.load(states[i])
                prop.load_perturbation(lam)
                prop.solver.adj_step(prop.t0 + (start + i) * prop.dt, prop.dt)
                lam = prop.save_perturbation()
        return lam, None, None, None


def advance(prop, q0, nsteps, checkpoint_every=100):
    """Differentiable q(T) = Phi^nsteps(q0) with RK-step checkpointing every ``checkpoint_every`` steps."""
    return _Advance.apply(q0, prop, int(nsteps), max(1, int(checkpoint_every)))


def state_tensor(case, q):
    """Host field dict -> (5, nz, ny, nx) float64 CUDA tensor."""
    arr = np.stack([np.asarray(q[k], dtype=np.float64).reshape((case.nx, case.ny, case.nz), order="F").T
                    for k in "UVWRE"])
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


def optimal_energy_gradient(prop, q0, dq0, nsteps, checkpoint_every=100):
    """Gradient of E(T) = 1/2 ||q'(T)||^2 w.r.t. the initial perturbation q'(0) = dq0, through
    the linearised propagator about the trajectory from q0 (adjoint-based optimal perturbations)."""
    # tangent forward: q'(T) = J dq0, keeping the base checkpoints
    prop.load(q0)
    prop.load_perturbation(dq0)
    checkpoints = []
    for s in range(nsteps):
        if s % checkpoint_every == 0:
            checkpoints.append(prop.save())
        prop.solver.tlm_step(prop.t0 + s * prop.dt, prop.dt)
    qT = prop.save_perturbation()
    energy = 0.5 * float((qT * qT).sum().item())
    lam = qT.clone()
    for seg in reversed(range(len(checkpoints))):
        start = seg * checkpoint_every
        length = min(checkpoint_every, nsteps - start)
        prop.load(checkpoints[seg])
        states = [checkpoints[seg]]
        for i in range(1, length):
            prop.steps(1, prop.t0 + (start + i - 1) * prop.dt)
            states.append(prop.save())
        for i in reversed(range(length)):
            prop.load(states[i])
            prop.load_perturbation(lam)
            prop.solver.adj_step(prop.t0 + (start + i) * prop.dt, prop.dt)
            lam = prop.save_perturbation()
    return energy, lam
