"""Batched Jacobian-free linear stability analysis (calcInstability.m) on the GPU library.

The reference (``calcInstability.m:1-459``) builds an Arnoldi basis of the
time-T propagator of the DNS: for each Krylov vector it writes a perturbed
flow file, launches a full DNS process for T, reads the result back and
forms the finite difference ``B = (Phi(q0 + eps sqrt(N) zeta) - Phi(q0)) /
(eps sqrt(N))`` (first order, ``:385-388``; second order ``:370-385``),
followed by classical Gram-Schmidt (``:389-398``) and ``eig`` of the
Hessenberg matrix with ``lambda = log(mu)/dT`` (``:429-436``).

Here a block of ``p`` Krylov vectors is a block of ``p`` members of one library
context.  The members are perturbed on the device (``gatt_perturb``),
advanced together by ``gatt_advance`` (one captured HIP graph per double
step for the whole block) and collected back into Krylov vectors
(``gatt_collect``).  Block Gram-Schmidt runs on the device
(``gatt_block_dot`` / ``gatt_block_update``), with a second
orthogonalisation pass and Cholesky-QR of each new block.  Only the small
projected matrices cross to the host, where numpy does the Cholesky and
eigenvalue factorizations (``eig`` in the reference).

Block Arnoldi with block size 1 and a single Gram-Schmidt pass is the
reference's scalar algorithm.
"""
import math
import os

import numpy as np

from .runtime import GattError, Solver


# ----------------------------------------------------------------- packing
def flow_region_indices(case, remove_buffer=False, buffers=None, single_beta=False):
    """0-based column-major indices of the flow-region points (calcInstability.m:276-281).

    ``case.flow_region`` is True at flow points (the reference takes
    ``~isnan(baseflow.U)``, i.e. everything outside walls).  With
    ``remove_buffer`` the ``buffers`` = ((xi, xf), (yi, yf), (zi, zf)) point
    counts at each end are removed as well.  With ``single_beta`` the indices are
    those of the z = 0 plane (``flowRegion(:,:,1)``, calcInstability.m:505, 555).
    """
    fr = np.ones((case.nx, case.ny, case.nz), dtype=bool) if case.flow_region is None else \
        np.array(case.flow_region, dtype=bool).reshape((case.nx, case.ny, case.nz))
    if single_beta:
        fr = fr[:, :, :1]
    if remove_buffer and buffers:
        fr = fr.copy()
        for d, (bi, bf) in enumerate(buffers):
            sl = [slice(None)] * 3
            if bi:
                sl[d] = slice(0, bi)
                fr[tuple(sl)] = False
            if bf:
                sl[d] = slice(fr.shape[d] - bf, None)
                fr[tuple(sl)] = False
    return np.flatnonzero(fr.ravel(order="F")).astype(np.int64)


def initial_disturbance(case, region, single_beta=False):
    """Gaussian bump at the finest mesh point in every field, normalised (calcInstability.m:691-741);
    with ``single_beta`` the bump lives in the x-y plane and fills all five slots (:726-737)."""
    X, Y, Z = (np.asarray(case.X, float), np.asarray(case.Y, float), np.asarray(case.Z, float))

    def centre(a):
        return a[int(np.argmin(np.gradient(a)))] if len(a) > 1 else a[0]

    ax = 10.0 / (X[-1] - X[0]) ** 2
    ay = 10.0 / (Y[-1] - Y[0]) ** 2
    arg = ax * (X[:, None, None] - centre(X)) ** 2 + ay * (Y[None, :, None] - centre(Y)) ** 2
    if single_beta:
        eta = np.exp(-arg).ravel(order="F")[region]
        z = np.concatenate([eta] * 5)
        return z / np.linalg.norm(z)
    if case.nz > 1:
        az = 10.0 / (Z[-1] - Z[0]) ** 2
        arg = arg + az * (Z[None, None, :] - centre(Z)) ** 2
    eta = np.exp(-arg).ravel(order="F")[region]
    nslots = 5 if case.nz > 1 else 4
    z = np.concatenate([eta] * nslots)
    return z / np.linalg.norm(z)


def vec2flow(case, region, vec, single_beta=False):
    """Host-side vec2flow (calcInstability.m:517-582): fields U..E with zeros outside the region.

    With ``single_beta`` the result is the x-y plane amplitude (nx, ny, 1) of the
    five fields, before the cos/sin z factors (the layout vec2modes writes).
    """
    shape = (case.nx, case.ny, 1 if single_beta else case.nz)
    np_ = shape[0] * shape[1] * shape[2]
    slots = "UVWRE" if case.nz > 1 else "UVRE"
    nf = len(region)
    out = {}
    for k in "UVWRE":
        a = np.zeros(np_)
        if k in slots:
            s = slots.index(k)
            a[region] = vec[s * nf:(s + 1) * nf]
        out[k] = a.reshape(shape, order="F")
    return out


# --------------------------------------------------------- block Arnoldi
class NumpyBlocks:
    """Host backend of :func:`block_arnoldi` (blocks are (columns, n) numpy arrays)."""

    def empty(self, cols, n):
        return np.zeros((cols, n))

    def dot(self, X, Y):
        return X @ Y.T

    def update(self, X, M, alpha, beta, Y):
        Y *= alpha
        Y += beta * (M.T @ X)

    def upload(self, host, dst):
        dst[...] = host

    def download(self, src):
        return np.array(src)


class DeviceBlocks:
    """Device backend: torch float64 tensors as storage, the library for arithmetic."""

    def __init__(self, solver):
        import torch
        self.torch = torch
        self.solver = solver

    def empty(self, cols, n):
        return self.torch.zeros((cols, n), dtype=self.torch.float64, device="cuda")

    def dot(self, X, Y):
        return self.solver.block_dot(X, Y)

    def update(self, X, M, alpha, beta, Y):
        self.solver.block_update(X, M, alpha, beta, Y)

    def upload(self, host, dst):
        self.solver.sync()
        dst.copy_(self.torch.from_numpy(np.ascontiguousarray(host)).to(dst.device))
        self.torch.cuda.synchronize()

    def download(self, src):
        self.solver.sync()
        return src.cpu().numpy()


def _chol_upper(G):
    """Upper R with G = R^T R, or None when G is not numerically positive definite."""
    try:
        L = np.linalg.cholesky(0.5 * (G + G.T))
    except np.linalg.LinAlgError:
        return None
    R = L.T
    d = np.abs(np.diag(R))
    if d.min() <= 1e-7 * d.max():
        return None
    return R


def orthonormalize(be, W, T, rng=None):
    """Cholesky-QR twice: W (p, n) -> orthonormal columns, written back into W; T is a work block.

    Returns (R, nk): the p x p matrix with W_in = Q R and the number of
    directions kept.  A rank-deficient block (nk < p) is completed with
    random directions orthogonal to the kept ones (their rows of R are zero).
    """
    p = W.shape[0]
    G = be.dot(W, W)
    R1 = _chol_upper(G)
    if R1 is None:
        # eigen-decomposition: keep the well-conditioned directions, refill the rest
        w, U = np.linalg.eigh(0.5 * (G + G.T))
        keep = w > 1e-14 * max(w.max(), 1e-300)
        M = np.zeros((p, p))
        M[:, keep] = U[:, keep] / np.sqrt(w[keep])
        be.update(W, M, 0.0, 1.0, T)
        R1 = np.zeros((p, p))
        R1[keep, :] = (U[:, keep] * np.sqrt(w[keep])).T
        nk = int(keep.sum())
        # orthonormal columns first, then random fill orthogonalised against them
        order = np.concatenate([np.flatnonzero(keep), np.flatnonzero(~keep)])
        be.update(T, np.eye(p)[:, order], 0.0, 1.0, W)
        R1 = R1[order, :]
        if nk < p:
            rng = rng or np.random.Generator(np.random.PCG64(12345))
            host = be.download(W)
            host[nk:] = rng.standard_normal((p - nk, host.shape[1]))
            for i in range(nk, p):
                for _ in range(2):
                    host[i] -= host[:i].T @ (host[:i] @ host[i])
                host[i] /= np.linalg.norm(host[i])
            be.upload(host, W)
        return R1, nk
    be.update(W, np.linalg.inv(R1), 0.0, 1.0, T)
    G2 = be.dot(T, T)
    R2 = _chol_upper(G2)
    if R2 is None:
        R2 = np.eye(p)
    be.update(T, np.linalg.inv(R2), 0.0, 1.0, W)
    return R2 @ R1, p


def block_arnoldi(apply, X0, m, be, passes=2, log=None, checkpoint=None, save_every=1):
    """Block Arnoldi of the operator ``apply(V, W)`` (W = A V for blocks of p columns).

    ``X0`` is the (p, n) starting block (on the backend), ``m`` the Krylov
    dimension (a multiple of p).  Returns (V, H, R0): the basis block (m + p,
    n), the (m + p) x m block Hessenberg matrix and the R of the starting
    block.  With p = 1 and passes = 1 this is the loop of calcInstability.m:
    384-398 (classical Gram-Schmidt).

    ``checkpoint``: a .npz file holding the basis, H and the number of finished
    blocks, written every ``save_every`` blocks and at the end; when it exists
    the iteration resumes from it (calcInstability.m:297-310, 406-423: the
    results file with zeta, H and M), with X0 ignored.  A larger m continues the
    saved run.
    """
    p, n = X0.shape
    if m % p:
        raise ValueError("Krylov dimension must be a multiple of the block size")
    nb = m // p
    V = be.empty(m + p, n)
    T = be.empty(p, n)
    H = np.zeros((m + p, m))
    j0 = 0
    if checkpoint is not None and os.path.exists(checkpoint):
        with np.load(checkpoint, allow_pickle=False) as d:
            if int(d["p"]) != p or d["V"].shape[1] != n:
                raise ValueError("checkpoint %s holds blocks of another shape" % checkpoint)
            j0 = min(int(d["blocks"]), nb)
            be.upload(np.ascontiguousarray(d["V"][:(j0 + 1) * p]), V[0:(j0 + 1) * p])
            H[:(j0 + 1) * p, :j0 * p] = d["H"][:(j0 + 1) * p, :j0 * p]
            R0 = np.array(d["R0"])
        if log:
            log("resuming from %s after %d blocks" % (checkpoint, j0))
    else:
        be.update(X0, np.eye(p), 0.0, 1.0, V[0:p])
        R0, _ = orthonormalize(be, V[0:p], T)

    def save(blocks):
        tmp = checkpoint + ".tmp.npz"
        np.savez(tmp, V=be.download(V[0:(blocks + 1) * p]), H=H[:(blocks + 1) * p, :blocks * p],
                 blocks=blocks, p=p, R0=R0)
        os.replace(tmp, checkpoint)

    for j in range(j0, nb):
        cur = V[j * p:(j + 1) * p]
        nxt = V[(j + 1) * p:(j + 2) * p]
        apply(cur, nxt)
        basis = V[0:(j + 1) * p]
        for _ in range(passes):
            h = be.dot(basis, nxt)
            be.update(basis, h, 1.0, -1.0, nxt)
            H[0:(j + 1) * p, j * p:(j + 1) * p] += h
        R, nk = orthonormalize(be, nxt, T)
        if nk < p:
            # exhausted directions were refilled at random: make them orthogonal to the basis
            # too (the kept directions already are, so their part of R is unchanged)
            for _ in range(2):
                be.update(basis, be.dot(basis, nxt), 1.0, -1.0, nxt)
            R2, _ = orthonormalize(be, nxt, T)
            R = R2 @ R
        H[(j + 1) * p:(j + 2) * p, j * p:(j + 1) * p] = R
        if log:
            log("block %d/%d: subdiagonal norm %.3e" % (j + 1, nb, np.linalg.norm(
                H[(j + 1) * p:(j + 2) * p, j * p:(j + 1) * p])))
        if checkpoint is not None and ((j + 1) % save_every == 0 or j + 1 == nb):
            save(j + 1)
    return V, H, R0


def ritz(H, m, T):
    """Eigenvalues mu of H(1:m,1:m), lambda = log(mu)/T sorted by real part (calcInstability.m:429-436)."""
    mu, y = np.linalg.eig(H[:m, :m])
    lam = np.log(mu.astype(complex)) / T
    order = np.argsort(-lam.real, kind="stable")
    return lam[order], mu[order], y[:, order]


# ------------------------------------------------------------- the driver
class Stability:
    """Time-T propagator of one case, applied to blocks of Krylov vectors on the GPU.

    ``block`` vectors are advanced per launch sequence (``2*block`` members with
    second-order differences).  ``epsilon`` is the reference's disturbance
    magnitude: perturbations are ``epsilon*sqrt(N)*zeta`` for unit vectors
    zeta (calcInstability.m:345).
    """

    def __init__(self, case, base, T, block=32, epsilon=1e-7, first_order=True, region=None, tune=False,
                 single_beta=False):
        self.case = case
        self.single_beta = bool(single_beta)
        self.T = float(T)
        self.nsteps = max(1, int(math.ceil(self.T / case.dt - 1e-9)))
        self.dt = self.T / self.nsteps
        self.block = block
        self.first_order = first_order
        self.region = flow_region_indices(case, single_beta=self.single_beta) if region is None else \
            np.asarray(region, dtype=np.int64)
        self.solver = Solver(case, batch=block if first_order else 2 * block)
        if tune:
            from .autotune import tune as _tune
            _tune(self.solver)
        if self.single_beta:
            self.solver.set_single_beta(True)
        self.solver.set_flow_region(self.region)
        self.solver.set_base(base)
        self.n = self.solver.vector_length()
        self.scale = epsilon * math.sqrt(self.n)
        self.be = DeviceBlocks(self.solver)
        self.ref = None
        self.advanced_members = 0
        if first_order:
            # UM = flow2vec(Phi(q0)) once (calcInstability.m:323-336)
            self.ref = self.be.empty(1, self.n)
            s = self.solver
            s.perturb(0, s.batch, None, 0.0)
            self._run()
            s.collect(0, 1, self.ref, 1.0)

    def _run(self):
        s = self.solver
        s.apply_bound_cond(0.0)                 # main.F90:112 on the perturbed initial flow
        s.advance(self.nsteps, 0.0, self.dt)
        self.advanced_members += s.batch

    def apply(self, V, W):
        """W = (Phi(q0 + s V) - Phi(q0)) / s column by column (s = epsilon sqrt(N)), for p <= block."""
        p = V.shape[0]
        if p > self.block:
            raise GattError("block larger than the context")
        s = self.solver
        if self.first_order:
            s.perturb(0, p, V, self.scale)
            if p < s.batch:
                s.perturb(p, s.batch - p, None, 0.0)
            self._run()
            s.collect(0, p, W, 1.0 / self.scale, ref=self.ref)
        else:
            s.perturb(0, p, V, self.scale)
            s.perturb(self.block, p, V, -self.scale)
            if p < self.block:
                s.perturb(p, self.block - p, None, 0.0)
                s.perturb(self.block + p, self.block - p, None, 0.0)
            self._run()
            s.collect(0, p, W, 0.5 / self.scale, minus_first=self.block)

    def close(self):
        self.solver.close()


def calc_instability(case, base, T, krylov_dim=64, block=32, epsilon=1e-7, first_order=True, seed=20260,
                     start="random", n_modes=0, region=None, log=None, tune=False, single_beta=False,
                     checkpoint=None, save_every=1):
    """Leading eigenvalues (and optionally modes) of the time-T propagator about ``base``.

    start = "random": ``block`` seeded N(0,1) vectors (numpy PCG64 ``seed``);
    start = "bump": calcInitialDisturb's Gaussian in the first column, random
    in the others.  Returns a dict with lambda, mu, H, the Krylov dimension, the
    propagator's step count and dt, the flow region (point indices), and ``modes`` (list of field dicts of the
    real and imaginary parts of the first ``n_modes`` Ritz vectors,
    vec2modes of calcInstability.m:437-446).  ``single_beta`` selects the
    singleBeta packing (one spanwise wavenumber, the z mesh one period of it;
    the reference's default for 3-D bases, calcInstability.m:14).  ``checkpoint`` /
    ``save_every``: the restart file of :func:`block_arnoldi`.
    """
    st = Stability(case, base, T, block=block, epsilon=epsilon, first_order=first_order, region=region,
                   tune=tune, single_beta=single_beta)
    try:
        rng = np.random.Generator(np.random.PCG64(seed))
        X0h = rng.standard_normal((block, st.n))
        if start == "bump":
            X0h[0] = initial_disturbance(case, st.region, single_beta)
        X0 = st.be.empty(block, st.n)
        st.be.upload(X0h, X0)
        V, H, _ = block_arnoldi(st.apply, X0, krylov_dim, st.be, log=log, checkpoint=checkpoint,
                                save_every=save_every)
        lam, mu, y = ritz(H, krylov_dim, st.T)
        out = {"lambda": lam, "mu": mu, "H": H, "m": krylov_dim, "nsteps": st.nsteps, "dt": st.dt,
               "members_advanced": st.advanced_members, "region": np.asarray(st.region)}
        if n_modes:
            k = min(n_modes, krylov_dim)
            basis = V[0:krylov_dim]
            re = st.be.empty(k, st.n)
            im = st.be.empty(k, st.n)
            st.be.update(basis, np.ascontiguousarray(y[:, :k].real), 0.0, 1.0, re)
            st.be.update(basis, np.ascontiguousarray(y[:, :k].imag), 0.0, 1.0, im)
            reh, imh = st.be.download(re), st.be.download(im)
            out["modes"] = [(vec2flow(case, st.region, reh[i], single_beta),
                              vec2flow(case, st.region, imh[i], single_beta)) for i in range(k)]
        return out
    finally:
        st.close()


# ------------------------------------------------------------ output files
def write_results(path, V, H, m):
    """The results file of calcInstability.m:419 (MAT 7.3): zeta (the n x m Krylov basis, one
    column per vector), H (the Hessenberg matrix) and M (the Krylov dimension)."""
    from . import flowio
    flowio.write_mat(path, {"zeta": np.asarray(V[:m]).T, "H": np.asarray(H), "M": float(m)})


def modes_fields(case, region, out, single_beta=False):
    """vec2modes (calcInstability.m:584-688) of the Ritz vectors in ``out["modes"]``: mode i is
    divided by its largest modulus and placed on the flow region, NaN elsewhere (W stays NaN in
    2-D); a mode whose eigenvalue has a negative imaginary part stays NaN (the conjugate of
    another).  Returns U..E of size (nx, ny, nz or 1, k), complex."""
    k = len(out["modes"])
    nz = 1 if (single_beta or case.nz == 1) else case.nz
    shape = (case.nx, case.ny, nz)
    fields = {f: np.full(shape + (k,), np.nan + 0j, order="F") for f in "UVWRE"}
    used = "UVRE" if case.nz == 1 else "UVWRE"
    for i, (re, im) in enumerate(out["modes"]):
        if out["lambda"][i].imag < 0:
            continue
        flat = {f: (re[f] + 1j * im[f]).reshape(-1, order="F")[region] for f in used}
        scale = max(np.max(np.abs(flat[f])) for f in used)
        for f in used:
            a = np.full(shape[0] * shape[1] * shape[2], np.nan + 0j)
            a[region] = flat[f] / scale
            fields[f][..., i] = a.reshape(shape, order="F")
    return fields


def write_modes(path, case, region, out, single_beta=False):
    """The modes file of calcInstability.m:437-452 (MAT 7.3): U, V, W, R, E of
    :func:`modes_fields`, lambda (every Ritz value, sorted by real part), X, Y, Z and nModes."""
    from . import flowio
    f = modes_fields(case, region, out, single_beta)
    f.update({"lambda": np.asarray(out["lambda"]), "X": np.asarray(case.X, float), "Y": np.asarray(case.Y, float),
              "Z": np.asarray(case.Z, float), "nModes": float(len(out["modes"]))})
    flowio.write_mat(path, f)
