"""ctypes wrapper of the CPU oracle (``oracle/gatt_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from gatt_dns_amd.abi import DP, build_desc, field_ptrs

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")


def build(force=False):
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        V = C.c_void_p
        L.oracle_create.restype = V
        L.oracle_create.argtypes = [V]
        L.oracle_destroy.argtypes = [V]
        L.oracle_solve.argtypes = [V, C.c_int, C.c_int, DP, DP]
        L.oracle_apply_filters.argtypes = [V, C.c_double, C.POINTER(DP)]
        L.oracle_nseq.argtypes = [V, C.POINTER(DP), C.POINTER(DP)]
        L.oracle_apply_bound_cond.argtypes = [V, C.c_double, C.POINTER(DP)]
        L.oracle_rk4.argtypes = [V, C.c_double, C.c_double, C.POINTER(DP), C.POINTER(DP)]
        L.oracle_euler.argtypes = [V, C.c_double, C.c_double, C.POINTER(DP), C.POINTER(DP)]
        L.oracle_ssprk3.argtypes = [V, C.c_double, C.c_double, C.POINTER(DP), C.POINTER(DP)]
        L.oracle_time_step.argtypes = [V, C.c_double, C.c_double, C.POINTER(DP), C.POINTER(DP)]
        L.oracle_init_mean.argtypes = [V, C.POINTER(DP)]
        L.oracle_get_mean.argtypes = [V, C.POINTER(DP)]
        L.oracle_sfd.argtypes = [V, C.c_double, C.POINTER(DP)]
        L.oracle_step.argtypes = [V, C.c_double, C.c_double, C.POINTER(DP), DP, DP]
        L.oracle_step.restype = C.c_int
        _lib = L
    return _lib


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class Oracle:
    """CPU restatement of the reference path for one :class:`Case`."""

    def __init__(self, case):
        self.case = case
        self.desc = build_desc(case)
        self.lib = lib()
        self.ctx = self.lib.oracle_create(C.cast(self.desc.ptr, C.c_void_p))
        self.shape = (case.nx, case.ny, case.nz)

    def __del__(self):
        try:
            self.lib.oracle_destroy(self.ctx)
        except Exception:
            pass

    def _state(self, q):
        return {k: _f(q[k]).reshape(self.shape, order="F").copy(order="F") for k in "UVWRE"}

    def _ptrs(self, st):
        return field_ptrs([st[k] for k in "UVWRE"])

    def solve(self, direction, kind, F):
        F = _f(F).reshape(self.shape, order="F")
        out = np.zeros(self.shape, order="F")
        self.lib.oracle_solve(self.desc.ptr, direction, kind, F.ctypes.data_as(DP), out.ctypes.data_as(DP))
        return out

    def rhs(self, q):
        st = self._state(q)
        d = {k: np.zeros(self.shape, order="F") for k in "UVWRE"}
        self.lib.oracle_nseq(self.ctx, self._ptrs(st), self._ptrs(d))
        return d

    def apply_bound_cond(self, q, t):
        st = self._state(q)
        self.lib.oracle_apply_bound_cond(self.ctx, t, self._ptrs(st))
        return st

    def apply_filters(self, q, theta):
        st = self._state(q)
        self.lib.oracle_apply_filters(self.ctx, theta, self._ptrs(st))
        return st

    def rk4(self, q, t, dt):
        st = self._state(q)
        out = {k: np.zeros(self.shape, order="F") for k in "UVWRE"}
        self.lib.oracle_rk4(self.ctx, t, dt, self._ptrs(st), self._ptrs(out))
        if self.case.nz == 1:
            out["W"] = st["W"]
        return out

    def _integrate(self, fn, q, t, dt):
        st = self._state(q)
        out = {k: np.zeros(self.shape, order="F") for k in "UVWRE"}
        fn(self.ctx, t, dt, self._ptrs(st), self._ptrs(out))
        if self.case.nz == 1:
            out["W"] = st["W"]
        return out

    def euler(self, q, t, dt):
        """EULER_2D / EULER_3D (timeIntegration.F90:158-222, 500-566)."""
        return self._integrate(self.lib.oracle_euler, q, t, dt)

    def ssprk3(self, q, t, dt):
        """SSPRK3_2D / SSPRK3_3D (timeIntegration.F90:225-335, 569-685)."""
        return self._integrate(self.lib.oracle_ssprk3, q, t, dt)

    def time_step(self, q, t, dt):
        """The integrator the case's time_stepping selects (main.F90:255-350)."""
        return self._integrate(self.lib.oracle_time_step, q, t, dt)

    def init_mean(self, q):
        st = self._state(q)
        self.lib.oracle_init_mean(self.ctx, self._ptrs(st))

    def mean(self):
        out = {k: np.zeros(self.shape, order="F") for k in "UVWRE"}
        self.lib.oracle_get_mean(self.ctx, self._ptrs(out))
        return out

    def sfd(self, q, dt):
        st = self._state(q)
        self.lib.oracle_sfd(self.ctx, dt, self._ptrs(st))
        return st

    def step(self, q, t, dt):
        """One main-loop step; returns (new state, umax[3], maxchange[5], diverged)."""
        st = self._state(q)
        um = np.zeros(3)
        mc = np.zeros(5)
        dv = self.lib.oracle_step(self.ctx, t, dt, self._ptrs(st), um.ctypes.data_as(DP), mc.ctypes.data_as(DP))
        return st, um, mc, bool(dv)

    def run(self, q, t0, dt, nsteps):
        st = self._state(q)
        um = np.zeros(3)
        mc = np.zeros(5)
        ptrs = self._ptrs(st)
        dv = 0
        t = t0
        for s in range(nsteps):
            dv = self.lib.oracle_step(self.ctx, t, dt, ptrs, um.ctypes.data_as(DP), mc.ctypes.data_as(DP))
            if dv:
                break
            t = t + dt   # main.F90:424
        return st, mc, bool(dv)
