"""Python binding of libgatt_dns (ctypes over the C ABI in include/gatt_dns.h).

:class:`Solver` mirrors the reference's subroutine boundary: ``rk4`` is
``RK4_3D``/``RK4_2D``, ``apply_filters`` is ``applyFilters3D``/``2D``,
``apply_bound_cond`` is ``applyBoundCond``, ``rhs`` is ``nseq3d``/``nseq2d``
and ``step`` is one iteration of the time loop of ``main.F90``.  The state
lives on the GPU; arrays cross the boundary only in ``set_state``/``get_state``.

There is no CPU fallback: constructing a Solver without the HIP library or a
GPU raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import DP, build_desc, field_ptrs

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgatt_dns.so")

EXPORTS = ["gatt_last_error", "gatt_device_info", "gatt_create", "gatt_destroy", "gatt_set_state",
           "gatt_get_state", "gatt_set_mean", "gatt_get_mean", "gatt_rk4", "gatt_apply_filters",
           "gatt_apply_bound_cond", "gatt_sfd", "gatt_rhs", "gatt_step", "gatt_advance",
           "gatt_line_solve", "gatt_sync", "gatt_time_kernel", "gatt_set_line_shape",
           "gatt_get_line_shape", "gatt_set_flow_region", "gatt_vector_length", "gatt_set_base",
           "gatt_perturb", "gatt_collect", "gatt_block_dot", "gatt_block_update", "gatt_group_create",
           "gatt_group_destroy", "gatt_group_advance", "gatt_set_perturbation", "gatt_get_perturbation",
           "gatt_copy_fields", "gatt_tlm_step", "gatt_adj_step", "gatt_line_solve_t", "gatt_euler",
           "gatt_ssprk3", "gatt_set_single_beta", "gatt_get_points", "gatt_get_previous_state",
           "gatt_checkpoint", "gatt_block_dot_noalloc"]

SLOT_STATE = 0
SLOT_PERTURBATION = 1
SLOT_MEAN = 2
SLOT_MEAN_PERTURBATION = 3


class GattError(RuntimeError):
    pass


def _after_torch(*tensors):
    """Wait for torch's pending work on the tensors' device before the library reads or writes them.

    The library runs on its own non-blocking stream, which is not ordered after torch's current
    stream (a fill or copy torch has queued on a tensor may still be running)."""
    for t in tensors:
        if t is not None:
            import torch
            torch.cuda.current_stream(t.device).synchronize()
            return


_lib = None


def load_library(path=LIB_PATH):
    """Load libgatt_dns.so and declare its signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GattError("libgatt_dns.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    V = C.c_void_p
    I = C.c_int32
    L.gatt_last_error.restype = C.c_char_p
    L.gatt_device_info.argtypes = [C.c_char_p, I]
    L.gatt_create.argtypes = [V, I, C.POINTER(V)]
    L.gatt_destroy.argtypes = [V]
    for name in ("gatt_set_state", "gatt_get_state", "gatt_set_mean", "gatt_get_mean", "gatt_get_previous_state"):
        getattr(L, name).argtypes = [V, I, C.POINTER(DP)]
    L.gatt_get_points.argtypes = [V, I, I, V, V]
    L.gatt_rk4.argtypes = [V, C.c_double, C.c_double]
    L.gatt_euler.argtypes = [V, C.c_double, C.c_double]
    L.gatt_set_single_beta.argtypes = [V, C.c_int32]
    L.gatt_ssprk3.argtypes = [V, C.c_double, C.c_double]
    L.gatt_apply_filters.argtypes = [V, C.c_double]
    L.gatt_apply_bound_cond.argtypes = [V, C.c_double]
    L.gatt_sfd.argtypes = [V, C.c_double]
    L.gatt_rhs.argtypes = [V, C.POINTER(DP)]
    L.gatt_step.argtypes = [V, C.c_double, C.c_double, DP, DP, C.POINTER(I)]
    L.gatt_advance.argtypes = [V, I, C.c_double, C.c_double, DP, C.POINTER(I)]
    L.gatt_line_solve.argtypes = [V, I, I, I, C.POINTER(DP), C.POINTER(DP)]
    L.gatt_sync.argtypes = [V]
    L.gatt_checkpoint.argtypes = [V, I]
    L.gatt_time_kernel.argtypes = [V, I, I, DP, DP]
    L.gatt_set_line_shape.argtypes = [V, I, I, I, I, I]
    IPt = C.POINTER(I)
    L.gatt_get_line_shape.argtypes = [V, I, I, IPt, IPt, IPt, IPt]
    I64 = C.c_int64
    L.gatt_set_flow_region.argtypes = [V, I64, C.POINTER(I64)]
    L.gatt_vector_length.argtypes = [V, C.POINTER(I64)]
    L.gatt_set_base.argtypes = [V, C.POINTER(DP)]
    L.gatt_perturb.argtypes = [V, I, I, V, I64, C.c_double]
    L.gatt_collect.argtypes = [V, I, I, I, V, C.c_double, V, I64]
    L.gatt_block_dot.argtypes = [V, I64, I, I, V, I64, V, I64, DP]
    L.gatt_block_dot_noalloc.argtypes = [V, I64, I, I, V, I64, V, I64, DP]
    L.gatt_block_update.argtypes = [V, I64, I, I, V, I64, DP, C.c_double, C.c_double, V, I64]
    L.gatt_group_create.argtypes = [C.POINTER(V), I, I, C.POINTER(V)]
    L.gatt_group_destroy.argtypes = [V]
    L.gatt_group_advance.argtypes = [V, I, DP, DP, DP, C.POINTER(I)]
    L.gatt_set_perturbation.argtypes = [V, I, C.POINTER(DP)]
    L.gatt_get_perturbation.argtypes = [V, I, C.POINTER(DP)]
    L.gatt_copy_fields.argtypes = [V, I, I, I, V]
    L.gatt_tlm_step.argtypes = [V, C.c_double, C.c_double, DP, C.POINTER(I)]
    L.gatt_adj_step.argtypes = [V, C.c_double, C.c_double]
    L.gatt_line_solve_t.argtypes = [V, I, I, I, C.POINTER(DP), C.POINTER(DP)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise GattError("gatt error %d: %s" % (rc, _lib.gatt_last_error().decode()))


def device_info():
    L = load_library()
    buf = C.create_string_buffer(256)
    _check(L.gatt_device_info(buf, 256))
    return buf.value.decode()


class Solver:
    """Device-resident time advance of ``batch`` members of one case."""

    def __init__(self, case, batch=1):
        self.lib = load_library()
        self.case = case
        self.batch = batch
        self.shape = (case.nx, case.ny, case.nz)
        self.desc = build_desc(case)
        h = C.c_void_p()
        _check(self.lib.gatt_create(C.cast(self.desc.ptr, C.c_void_p), batch, C.byref(h)))
        self.ctx = h

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.gatt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def _arrays(self, q):
        return [np.asfortranarray(np.asarray(q[k], dtype=np.float64).reshape(self.shape, order="F"))
                for k in "UVWRE"]

    def set_state(self, q, member=0):
        arrs = self._arrays(q)
        _check(self.lib.gatt_set_state(self.ctx, member, field_ptrs(arrs)))

    def get_state(self, member=0):
        out = [np.zeros(self.shape, order="F") for _ in range(5)]
        _check(self.lib.gatt_get_state(self.ctx, member, field_ptrs(out)))
        return dict(zip("UVWRE", out))

    def get_points(self, points, member=0):
        """U, V, W, R, E of the current state at grid points (rows of 0-based (i, j, k), the
        reference's indTracked minus one): an (n, 5) array, read on the device."""
        ijk = np.asarray(points, dtype=np.int64).reshape(-1, 3)
        nx, ny, nz = self.shape
        if ijk.size and (np.any(ijk < 0) or np.any(ijk >= np.array([nx, ny, nz]))):
            raise GattError("tracked point outside the grid")
        idx = np.ascontiguousarray(ijk[:, 0] + nx * (ijk[:, 1] + ny * ijk[:, 2]))
        out = np.zeros((len(idx), 5))
        _check(self.lib.gatt_get_points(self.ctx, member, len(idx), idx.ctypes.data_as(C.c_void_p),
                                        out.ctypes.data_as(C.c_void_p)))
        return out

    def set_mean(self, q, member=0):
        _check(self.lib.gatt_set_mean(self.ctx, member, field_ptrs(self._arrays(q))))

    def get_mean(self, member=0):
        out = [np.zeros(self.shape, order="F") for _ in range(5)]
        _check(self.lib.gatt_get_mean(self.ctx, member, field_ptrs(out)))
        return dict(zip("UVWRE", out))

    # -- reference subroutines -------------------------------------------------
    def rk4(self, t, dt):
        _check(self.lib.gatt_rk4(self.ctx, t, dt))

    def euler(self, t, dt):
        _check(self.lib.gatt_euler(self.ctx, t, dt))

    def ssprk3(self, t, dt):
        _check(self.lib.gatt_ssprk3(self.ctx, t, dt))

    def apply_filters(self, theta):
        _check(self.lib.gatt_apply_filters(self.ctx, theta))

    def apply_bound_cond(self, t):
        _check(self.lib.gatt_apply_bound_cond(self.ctx, t))

    def sfd(self, dt):
        _check(self.lib.gatt_sfd(self.ctx, dt))

    def rhs(self):
        out = [np.zeros(self.shape, order="F") for _ in range(5)]
        _check(self.lib.gatt_rhs(self.ctx, field_ptrs(out)))
        return dict(zip("UVWRE", out))

    def step(self, t, dt):
        um = np.zeros(3)
        mc = np.zeros(5)
        dv = C.c_int32(0)
        _check(self.lib.gatt_step(self.ctx, t, dt, um.ctypes.data_as(DP), mc.ctypes.data_as(DP), C.byref(dv)))
        return um, mc, bool(dv.value)

    def advance(self, nsteps, t0, dt):
        mc = np.zeros(5)
        dv = C.c_int32(0)
        _check(self.lib.gatt_advance(self.ctx, nsteps, t0, dt, mc.ctypes.data_as(DP), C.byref(dv)))
        return mc, bool(dv.value)

    def get_previous_state(self, member=0):
        """The state before the last step (the pre-step U of a diverged step)."""
        out = [np.zeros(self.shape, order="F") for _ in range(5)]
        _check(self.lib.gatt_get_previous_state(self.ctx, member, field_ptrs(out)))
        return dict(zip("UVWRE", out))

    def checkpoint(self, op):
        """op 0: save the state (and SFD mean) of every member on the device; op 1: restore it."""
        _check(self.lib.gatt_checkpoint(self.ctx, op))

    def sync(self):
        _check(self.lib.gatt_sync(self.ctx))

    def time_kernel(self, kernel, reps):
        """(mean ms per launch, algorithmic bytes per launch) of one kernel class."""
        ms = C.c_double(0)
        by = C.c_double(0)
        _check(self.lib.gatt_time_kernel(self.ctx, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_line_shape(self, direction, kind, variant, param, waves_per_block=4):
        """variant 0: param = lanes per line; variants 1 and 2: param = points per lane."""
        _check(self.lib.gatt_set_line_shape(self.ctx, direction, kind, variant, param, waves_per_block))

    def get_line_shape(self, direction, kind):
        """(variant, G, M, waves per block) of the current shape."""
        v, g, m, w = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.lib.gatt_get_line_shape(self.ctx, direction, kind, C.byref(v), C.byref(g), C.byref(m),
                                            C.byref(w)))
        return v.value, g.value, m.value, w.value

    # -- stability analysis (device blocks are torch float64 tensors of shape
    #    (columns, n), one Krylov vector per row) -------------------------------
    def set_flow_region(self, idx=None):
        """0-based column-major indices of the flow-region points (None: every point)."""
        if idx is None:
            _check(self.lib.gatt_set_flow_region(self.ctx, 0, None))
            return
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        _check(self.lib.gatt_set_flow_region(self.ctx, len(idx), idx.ctypes.data_as(C.POINTER(C.c_int64))))

    def set_single_beta(self, on=True):
        """singleBeta packing of the Krylov vectors (resets the flow region)."""
        _check(self.lib.gatt_set_single_beta(self.ctx, 1 if on else 0))

    def vector_length(self):
        n = C.c_int64(0)
        _check(self.lib.gatt_vector_length(self.ctx, C.byref(n)))
        return n.value

    def set_base(self, q):
        _check(self.lib.gatt_set_base(self.ctx, field_ptrs(self._arrays(q))))

    @staticmethod
    def _block(t):
        if t is None:
            return None, 0
        if str(t.dtype) != "torch.float64" or not t.is_cuda or t.stride(-1) != 1:
            raise GattError("device blocks must be float64 CUDA tensors with unit stride along a vector")
        _after_torch(t)
        return C.c_void_p(t.data_ptr()), (t.stride(0) if t.dim() == 2 else t.shape[-1])

    def _need(self, t, cols, what):
        if t is not None and (t.dim() != 2 or t.shape[0] < cols or t.shape[1] < self.vector_length()):
            raise GattError("%s must hold %d vectors of length %d" % (what, cols, self.vector_length()))

    def perturb(self, first, count, V, scale):
        """Members first..first+count-1 = base + scale * vec2flow(V[b]) (V None: base)."""
        self._need(V, count, "V")
        ptr, ld = self._block(V)
        _check(self.lib.gatt_perturb(self.ctx, first, count, ptr, ld, scale))

    def collect(self, first, count, W, scale, minus_first=-1, ref=None):
        """W[b] = scale * (flow2vec(member first+b) - flow2vec(member minus_first+b) or ref)."""
        self._need(W, count, "W")
        self._need(ref, 1, "ref")
        wp, ld = self._block(W)
        rp, _ = self._block(ref)
        _check(self.lib.gatt_collect(self.ctx, first, count, minus_first, rp, scale, wp, ld))

    def block_dot(self, X, Y, noalloc=False):
        """H = X^T Y (k x p numpy array) for blocks X (k, n) and Y (p, n).  ``noalloc`` calls
        gatt_block_dot_noalloc (no device allocation; blocks up to the static scratch limit)."""
        xp, ldx = self._block(X)
        yp, ldy = self._block(Y)
        k, p, n = X.shape[0], Y.shape[0], X.shape[1]
        if X.dim() != 2 or Y.dim() != 2 or Y.shape[1] != n:
            raise GattError("X and Y must be (columns, n) blocks of the same n")
        H = np.zeros((k, p), order="F")
        fn = self.lib.gatt_block_dot_noalloc if noalloc else self.lib.gatt_block_dot
        _check(fn(self.ctx, n, k, p, xp, ldx, yp, ldy, H.ctypes.data_as(DP)))
        return H

    def block_update(self, X, M, alpha, beta, Y):
        """Y = alpha Y + beta X M for blocks X (k, n), Y (p, n) and M (k x p)."""
        xp, ldx = self._block(X)
        yp, ldy = self._block(Y)
        M = np.asfortranarray(M, dtype=np.float64)
        k, p, n = X.shape[0], Y.shape[0], X.shape[1]
        if X.dim() != 2 or Y.dim() != 2 or Y.shape[1] != n:
            raise GattError("X and Y must be (columns, n) blocks of the same n")
        if M.shape != (k, p):
            raise GattError("M must be %d x %d" % (k, p))
        _check(self.lib.gatt_block_update(self.ctx, n, k, p, xp, ldx, M.ctypes.data_as(DP), alpha, beta, yp, ldy))

    # -- tangent and adjoint ---------------------------------------------------
    def set_perturbation(self, q, member=0):
        arrs = self._arrays(q)
        _check(self.lib.gatt_set_perturbation(self.ctx, member, field_ptrs(arrs)))

    def get_perturbation(self, member=0):
        out = [np.zeros(self.shape, order="F") for _ in range(5)]
        _check(self.lib.gatt_get_perturbation(self.ctx, member, field_ptrs(out)))
        return dict(zip("UVWRE", out))

    def copy_fields(self, slot, to_ctx, tensor, member=0):
        """Device copy of the state (slot 0), perturbation (1), SFD mean (2) or the mean's
        perturbation (3) to/from a float64 CUDA tensor of 5*nx*ny*nz contiguous values
        (U,V,W,R,E, each x fastest)."""
        npts = self.shape[0] * self.shape[1] * self.shape[2]
        if (str(tensor.dtype) != "torch.float64" or not tensor.is_cuda or not tensor.is_contiguous()
                or tensor.numel() != 5 * npts):
            raise GattError("expected a contiguous float64 CUDA tensor of %d values" % (5 * npts))
        _after_torch(tensor)
        _check(self.lib.gatt_copy_fields(self.ctx, member, slot, 1 if to_ctx else 0, C.c_void_p(tensor.data_ptr())))

    def tlm_step(self, t, dt):
        """State and tangent advance one step; returns (maxchange, diverged) of the state."""
        mc = np.zeros(5)
        dv = C.c_int32(0)
        _check(self.lib.gatt_tlm_step(self.ctx, t, dt, mc.ctypes.data_as(DP), C.byref(dv)))
        return mc, bool(dv.value)

    def adj_step(self, t, dt):
        """Perturbation <- J^T perturbation about the current state (the state is not advanced)."""
        _check(self.lib.gatt_adj_step(self.ctx, t, dt))

    def line_solve_t(self, direction, kind, fields_in, fields_out):
        """Transposed line solve (device tensors); fields_in are overwritten (workspace)."""
        n = len(fields_in)
        _after_torch(*fields_in, *fields_out)
        pin = (DP * n)(*[C.cast(C.c_void_p(t.data_ptr()), DP) for t in fields_in])
        pout = (DP * n)(*[C.cast(C.c_void_p(t.data_ptr()), DP) for t in fields_out])
        _check(self.lib.gatt_line_solve_t(self.ctx, direction, kind, n, pin, pout))

    def line_solve(self, direction, kind, fields_in, fields_out):
        """Device-pointer unit entry point: lists of torch CUDA tensors (float64, Fortran-ordered data)."""
        n = len(fields_in)
        _after_torch(*fields_in, *fields_out)
        pin = (DP * n)(*[C.cast(C.c_void_p(t.data_ptr()), DP) for t in fields_in])
        pout = (DP * n)(*[C.cast(C.c_void_p(t.data_ptr()), DP) for t in fields_out])
        _check(self.lib.gatt_line_solve(self.ctx, direction, kind, n, pin, pout))


class Group:
    """Ragged group of solvers (different cases) advanced together (gatt_group_*).

    mode 0 replays one HIP graph per double step with a branch per member;
    mode 1 replays per-member graphs round-robin on the members' streams;
    mode 2 replays one graph of ragged launches, each serving every member.
    """

    def __init__(self, solvers, mode=0):
        self.lib = load_library()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.ctx for s in self.solvers])
        h = C.c_void_p()
        _check(self.lib.gatt_group_create(arr, len(self.solvers), mode, C.byref(h)))
        self.handle = h

    def advance(self, nsteps, t0, dt):
        n = len(self.solvers)
        t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, float), (n,)))
        dt = np.ascontiguousarray(np.asarray(dt, float))
        mc = np.zeros((n, 5))
        dv = (C.c_int32 * n)()
        _check(self.lib.gatt_group_advance(self.handle, nsteps, t0.ctypes.data_as(DP), dt.ctypes.data_as(DP),
                                           mc.ctypes.data_as(DP), dv))
        return mc, [bool(v) for v in dv]

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gatt_group_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
