"""CPU checks of the stability-analysis pieces (calcInstability.m).

* The closed-form Fourier spectrum (tests/fourier_rest.py) is itself checked
  against the oracle's line solves on single modes.
* The oracle's Jacobian-free propagator (second-order differences) is pinned
  by that spectrum on a 12 x 12 periodic box at rest: every eigenvalue of the
  finite-difference Jacobian of the time-T map matches.
* The oracle's scalar Arnoldi (calcInstability.m:384-436) and the block
  Arnoldi of gatt_dns_amd/lst.py (host backend) are checked on matrices with
  known spectra; block size 1 with one Gram-Schmidt pass reproduces the
  scalar algorithm's Hessenberg matrix.
"""
import numpy as np
import pytest

import fourier_rest as FR
from gatt_dns_amd import lst
from gatt_dns_amd.prep import cases


@pytest.fixture(scope="module")
def rest12():
    return cases.periodic_rest(12, 12)


def test_fourier_symbols_match_line_solves(rest12, oracle_mod):
    o = oracle_mod.Oracle(rest12)
    n = 12
    x = np.arange(n) / n
    dsym, fsym = FR.derivative_symbol(n), FR.filter_symbol(n)
    for m in (1, 2, 5):
        F = np.sin(2 * np.pi * m * x)[:, None, None] * np.ones((n, n, 1))
        dF = o.solve(0, 0, F)[:, 0, 0]
        assert np.abs(dsym[m].real) < 1e-12
        np.testing.assert_allclose(dF, dsym[m].imag * np.cos(2 * np.pi * m * x), atol=1e-11)
        fF = o.solve(1, 1, np.transpose(F, (1, 0, 2)).copy(order="F"))[0, :, 0]
        np.testing.assert_allclose(fF, fsym[m].real * np.sin(2 * np.pi * m * x), atol=1e-13)
    # the filter removes the Nyquist mode and keeps constants
    assert abs(fsym[n // 2]) < 1e-13 and abs(fsym[0] - 1) < 1e-13


def test_oracle_propagator_spectrum_closed_form(rest12):
    from oracle import lst as OL
    T = 0.25
    P = OL.OraclePropagator(rest12, rest12.state0, T, epsilon=1e-5, first_order=False)
    N = P.N
    assert N == 4 * 144
    J = np.zeros((N, N))
    e = np.zeros(N)
    for i in range(N):
        e[:] = 0
        e[i] = 1
        J[:, i] = P.apply(e)
    mu = np.linalg.eigvals(J)
    mu_ref = FR.propagator_eigenvalues(rest12, P.nsteps, P.dt)
    d, r, c = FR.match(mu, mu_ref)
    assert d < 1e-6, d
    lead = np.abs(mu_ref[c]) > 0.5
    lam = np.log(mu[r][lead].astype(complex)) / T
    lam_ref = np.log(mu_ref[c][lead].astype(complex)) / T
    assert np.abs(lam - lam_ref).max() < 1e-6


def _known_matrix(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ev = np.concatenate([rng.uniform(-1, 1, n - 4), [1.5 + 0.5j, 1.5 - 0.5j, -1.2 + 0.3j, -1.2 - 0.3j]])
    S = rng.standard_normal((n, n)) + 3 * np.eye(n)
    D = np.zeros((n, n))
    i = 0
    while i < n:
        if abs(ev[i].imag) > 0 and i + 1 < n:
            D[i:i + 2, i:i + 2] = [[ev[i].real, ev[i].imag], [-ev[i].imag, ev[i].real]]
            i += 2
        else:
            D[i, i] = ev[i].real
            i += 1
    A = S @ D @ np.linalg.inv(S)
    return A, np.linalg.eigvals(A)


def test_oracle_scalar_arnoldi_known_matrix():
    from oracle import lst as OL
    A, ev = _known_matrix(30, 3)
    rng = np.random.Generator(np.random.PCG64(4))
    zeta, H = OL.arnoldi(lambda z: A @ z, rng.standard_normal(30), 30)
    d, _, _ = FR.match(np.linalg.eigvals(H), ev)
    assert d < 1e-6       # single-pass classical Gram-Schmidt over the full space
    # Arnoldi relation A zeta(:,1:M-1) = zeta H(:,1:M-1)
    np.testing.assert_allclose(A @ zeta[:, :-1], zeta @ H[:, :-1], atol=1e-10)


@pytest.mark.parametrize("p,passes", [(4, 2), (8, 2), (1, 2)])
def test_block_arnoldi_known_matrix(p, passes):
    A, ev = _known_matrix(48, 5)
    be = lst.NumpyBlocks()
    rng = np.random.Generator(np.random.PCG64(6))
    X0 = rng.standard_normal((p, 48))

    def apply(V, W):
        W[...] = V @ A.T

    V, H, _ = lst.block_arnoldi(apply, X0, 48, be, passes=passes)
    d, _, _ = FR.match(np.linalg.eigvals(H[:48, :48]), ev)
    assert d < 1e-8
    Q = V[:48]
    np.testing.assert_allclose(Q @ Q.T, np.eye(48), atol=1e-10)


def test_block_size_one_is_the_reference_loop():
    """Block size 1, one pass: the H of calcInstability.m:384-398 (up to rounding)."""
    from oracle import lst as OL
    A, _ = _known_matrix(20, 7)
    rng = np.random.Generator(np.random.PCG64(8))
    z0 = rng.standard_normal(20)
    _, Href = OL.arnoldi(lambda z: A @ z, z0, 12)

    def apply(V, W):
        W[...] = V @ A.T

    _, H, _ = lst.block_arnoldi(apply, z0[None, :].copy(), 12, lst.NumpyBlocks(), passes=1)
    # the reference leaves H(M+1,M) unset; signs of the subdiagonal agree (positive norms)
    np.testing.assert_allclose(H[:12, :12], Href, rtol=1e-9, atol=1e-11)


def test_flow2vec_packing_matches_oracle(oracle_mod):
    from oracle import lst as OL
    case = cases.lid_driven(17)
    region = lst.flow_region_indices(case)
    assert len(region) == 15 * 15               # walls on the four sides are not flow points
    rng = np.random.Generator(np.random.PCG64(9))
    v = rng.standard_normal(4 * len(region))
    q = lst.vec2flow(case, region, v)
    np.testing.assert_array_equal(OL.flow2vec(case, region, q), v)
    q2 = OL.vec2flow(case, region, v)
    for k in "UVWRE":
        np.testing.assert_array_equal(q[k], q2[k])
    z = lst.initial_disturbance(case, region)
    assert abs(np.linalg.norm(z) - 1) < 1e-14 and z.shape == v.shape


def test_block_arnoldi_breakdown_refill():
    """Krylov space exhausted early (6 distinct eigenvalues): refilled directions keep the basis orthonormal."""
    rng = np.random.Generator(np.random.PCG64(10))
    n, p, m = 48, 4, 32
    ev = np.repeat([0.9, 0.5, -0.3, 0.1, 1.2, -0.8], 8)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(ev) @ Q.T

    def apply(V, W):
        W[...] = V @ A.T

    V, H, _ = lst.block_arnoldi(apply, rng.standard_normal((p, n)), m, lst.NumpyBlocks())
    B = V[:m + p]
    np.testing.assert_allclose(B @ B.T, np.eye(m + p), atol=1e-10)
    ritz = np.linalg.eigvals(H[:m, :m])
    for e in np.unique(ev):
        assert np.abs(ritz - e).min() < 1e-10
    # the Arnoldi relation A V_m = V_{m+p} H holds through the refill
    np.testing.assert_allclose(V[:m] @ A.T, H.T @ B, atol=1e-10)


def test_single_beta_packing_round_trip(oracle_mod):
    """flow2vec(vec2flow(v)) = v for the singleBeta packing: (2/nz) sum_k cos^2(z_k) = 1 and
    the same for sin, so the projection weights invert the plane-k factors."""
    from oracle import lst as OL
    case = cases.periodic_rest(12, 14, 4)
    region = lst.flow_region_indices(case, single_beta=True)
    assert len(region) == 12 * 14
    rng = np.random.Generator(np.random.PCG64(13))
    v = rng.standard_normal(5 * len(region))
    q = OL.vec2flow(case, region, v, single_beta=True)
    np.testing.assert_allclose(OL.flow2vec(case, region, q, single_beta=True), v, rtol=0, atol=1e-15)
    # W carries sin(z): zero on the planes z = 0 and z = pi
    assert np.max(np.abs(q["W"][:, :, 0])) == 0.0 and np.max(np.abs(q["W"][:, :, 2])) < 1e-15
    plane = lst.vec2flow(case, region, v, single_beta=True)
    for k in "UVWRE":
        np.testing.assert_array_equal(plane[k][:, :, 0], q[k][:, :, 0] if k != "W" else q[k][:, :, 1])
    z = lst.initial_disturbance(case, region, single_beta=True)
    assert abs(np.linalg.norm(z) - 1) < 1e-14 and z.shape == v.shape


def test_single_beta_subspace_is_invariant(oracle_mod):
    """With the 4-point spectral z stencil (makeMatrices.m:28-38) the linearised step at rest
    maps U,V,R,E ~ cos(z), W ~ sin(z) into itself: the propagated disturbance is rebuilt
    from its singleBeta projection.  Pairing W with cos(z) instead would lose it."""
    from oracle import lst as OL
    case = cases.periodic_rest(12, 12, 4)
    assert case.deriv[2].nR == 2 and case.filt[2].nR == 1     # spectral z derivative, identity filter
    region = lst.flow_region_indices(case, single_beta=True)
    P = OL.OraclePropagator(case, case.state0, 4 * case.dt, epsilon=1e-6, first_order=False, single_beta=True)
    rng = np.random.Generator(np.random.PCG64(5))
    v = rng.standard_normal(P.N)
    v /= np.linalg.norm(v)
    s = P.eps * np.sqrt(P.N)
    qp, qm = P.phi(P.perturbed(v)), P.phi(P.perturbed(v, -1.0))
    D = {k: (qp[k] - qm[k]) / (2 * s) for k in "UVWRE"}
    rebuilt = OL.vec2flow(case, region, OL.flow2vec(case, region, D, single_beta=True), single_beta=True)
    scale = max(np.max(np.abs(D[k])) for k in "UVWRE")
    for k in "UVWRE":
        assert np.max(np.abs(rebuilt[k] - D[k])) <= 1e-6 * scale, k
    # the propagator did act: the result differs from the initial disturbance
    assert np.linalg.norm(P.apply(v) - v) > 1e-3


def test_oracle_single_beta_spectrum_closed_form(oracle_mod):
    """The oracle's singleBeta propagator on a 12 x 12 x 4 box at rest (spectral z): every
    eigenvalue of its finite-difference Jacobian is one of the x-y modes times the first z
    harmonic in closed form."""
    from oracle import lst as OL
    case = cases.periodic_rest(12, 12, 4)
    T = 0.05
    P = OL.OraclePropagator(case, case.state0, T, epsilon=1e-5, first_order=False, single_beta=True)
    N = P.N
    assert N == 5 * 144
    J = np.zeros((N, N))
    e = np.zeros(N)
    for i in range(N):
        e[:] = 0
        e[i] = 1
        J[:, i] = P.apply(e)
    mu = np.linalg.eigvals(J)
    mu_ref = FR.propagator_eigenvalues(case, P.nsteps, P.dt, single_beta=True)
    d, r, c = FR.match(mu, mu_ref)
    assert d < 1e-6, d


def test_block_arnoldi_restart(tmp_path):
    """The checkpoint file (calcInstability.m's results file): a run stopped after 3 blocks and
    resumed to 6 gives the basis and H of an uninterrupted 6-block run, bit for bit."""
    A, _ = _known_matrix(40, 7)
    be = lst.NumpyBlocks()
    X0 = np.random.Generator(np.random.PCG64(8)).standard_normal((4, 40))

    def apply(V, W):
        W[...] = V @ A.T

    Vf, Hf, Rf = lst.block_arnoldi(apply, X0, 24, be)
    ck = str(tmp_path / "results.npz")
    lst.block_arnoldi(apply, X0, 12, be, checkpoint=ck, save_every=2)
    calls = []

    def counted(V, W):
        calls.append(1)
        apply(V, W)

    Vr, Hr, Rr = lst.block_arnoldi(counted, np.zeros((4, 40)), 24, be, checkpoint=ck)
    assert len(calls) == 3                     # only the blocks after the saved ones were applied
    np.testing.assert_array_equal(Hr, Hf)
    np.testing.assert_array_equal(Vr, Vf)
    np.testing.assert_array_equal(Rr, Rf)
    with pytest.raises(ValueError):
        lst.block_arnoldi(apply, np.zeros((2, 40)), 24, be, checkpoint=ck)
