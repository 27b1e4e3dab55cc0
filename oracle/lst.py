"""CPU restatement of the reference's stability-analysis loop (calcInstability.m).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of
``gatt_dns_amd/lst.py``; the product never imports it.

* :func:`flow2vec` / :func:`vec2flow` follow calcInstability.m:480-515 and
  517-582 (2-D drops W); with ``single_beta`` the singleBeta branches
  (:493-509, :557-579): one x-y plane of amplitudes, cos(z) for U,V,R,E and
  sin(z) for W, projected with the weights (2/nz) cos(z_k), (2/nz) sin(z_k).
* :class:`OraclePropagator` is the reference's per-vector DNS launch
  (calcInstability.m:323-388): the perturbed flow gets the boundary
  conditions of main.F90:112 and is advanced nsteps fixed-dt steps by the
  oracle's main-loop step; first- or second-order differences.
* :func:`arnoldi` is the loop of calcInstability.m:384-420: classical
  Gram-Schmidt against all previous vectors (``H(j,k) = zeta(:,j)'*B`` with
  the unmodified B), ``H(k+1,k) = norm(uPrime)``.
* :func:`ritz` is calcInstability.m:429-436: ``eig(H(1:k,1:k))``,
  ``lambda = log(mu)/dT``, sorted by decreasing real part.
"""
import math

import numpy as np

from .oracle import Oracle


def slots(case):
    return "UVWRE" if case.nz > 1 else "UVRE"


def _zplanes(nz):
    return np.linspace(0.0, 2.0 * np.pi, nz + 1)[:-1]


def flow2vec(case, region, q, single_beta=False):
    if not single_beta:
        return np.concatenate([np.asarray(q[k], float).ravel(order="F")[region] for k in slots(case)])
    z = _zplanes(case.nz)
    out = []
    for k in "UVWRE":
        w = (np.sin(z) if k == "W" else np.cos(z)) * 2 / case.nz
        a = np.asarray(q[k], float).reshape((case.nx, case.ny, case.nz), order="F")
        acc = w[0] * a[:, :, 0]
        for kz in range(1, case.nz):      # sum(..., 3) in plane order
            acc = acc + w[kz] * a[:, :, kz]
        out.append(acc.ravel(order="F")[region])
    return np.concatenate(out)


def vec2flow(case, region, vec, single_beta=False):
    if single_beta:
        z = _zplanes(case.nz)
        nf = len(region)
        out = {}
        for s_, k in enumerate("UVWRE"):
            a = np.zeros(case.nx * case.ny)
            a[region] = vec[s_ * nf:(s_ + 1) * nf]
            w = np.sin(z) if k == "W" else np.cos(z)
            out[k] = a.reshape((case.nx, case.ny, 1), order="F") * w[None, None, :]
        return out
    shape = (case.nx, case.ny, case.nz)
    npts = case.nx * case.ny * case.nz
    nf = len(region)
    out = {}
    for k in "UVWRE":
        a = np.zeros(npts)
        if k in slots(case):
            s = slots(case).index(k)
            a[region] = vec[s * nf:(s + 1) * nf]
        out[k] = a.reshape(shape, order="F")
    return out


class OraclePropagator:
    def __init__(self, case, base, T, epsilon=1e-7, first_order=True, region=None, single_beta=False):
        self.case = case
        self.sb = single_beta
        self.base = {k: np.asarray(base[k], float) for k in "UVWRE"}
        self.nsteps = max(1, int(math.ceil(T / case.dt - 1e-9)))
        self.dt = T / self.nsteps
        npts = case.nx * case.ny * (1 if single_beta else case.nz)
        self.region = np.flatnonzero(np.ones(npts, bool)) if region is None else region
        self.N = len(slots(case)) * len(self.region)
        self.eps = epsilon
        self.first_order = first_order
        self.oracle = Oracle(case)
        self.UM = flow2vec(case, self.region, self.phi(self.base), single_beta) if first_order else None

    def phi(self, q):
        o = self.oracle
        st = o.apply_bound_cond(q, 0.0)
        if self.case.sfd_type > 0:
            o.init_mean(self.base)
        st, _, _ = o.run(st, 0.0, self.dt, self.nsteps)
        return st

    def perturbed(self, z, sign=1.0):
        d = vec2flow(self.case, self.region, z * (sign * self.eps * math.sqrt(self.N)), self.sb)
        return {k: self.base[k] + d[k] for k in "UVWRE"}

    def apply(self, z):
        s = self.eps * math.sqrt(self.N)
        UP = flow2vec(self.case, self.region, self.phi(self.perturbed(z)), self.sb)
        if self.first_order:
            return (UP - self.UM) / s
        UMm = flow2vec(self.case, self.region, self.phi(self.perturbed(z, -1.0)), self.sb)
        return (UP - UMm) / (2 * s)


def arnoldi(apply, z0, M):
    N = len(z0)
    zeta = np.zeros((N, M))
    H = np.zeros((M, M))
    zeta[:, 0] = z0 / np.linalg.norm(z0)
    for k in range(M):
        B = apply(zeta[:, k])
        u = B.copy()
        for j in range(k + 1):
            H[j, k] = zeta[:, j] @ B
            u = u - H[j, k] * zeta[:, j]
        if k < M - 1:
            H[k + 1, k] = np.linalg.norm(u)
            zeta[:, k + 1] = u / H[k + 1, k]
    return zeta, H


def ritz(H, k, dT):
    mu = np.linalg.eigvals(H[:k, :k])
    lam = np.log(mu.astype(complex)) / dT
    order = np.argsort(-lam.real, kind="stable")
    return lam[order], mu[order]
