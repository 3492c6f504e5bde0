"""CPU pinning of the oracle matvec (mv_AN/mv_N/mv_L) and Krylov solvers.

The FFT-circulant product is checked against the brute-force Toeplitz sum on
tiny grids (>= 5 per side, the n_medium region of assembly_nop.m), against the
adjoint identity, linearity and complex symmetry; pgmres against a dense
solve and against a Givens-rotation GMRES written independently here.
"""
import numpy as np
import pytest

from oracle import krylov, mvp
from oracle import operators as oop


def _parity_G(L, M, N, rng):
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    return G


@pytest.mark.parametrize("shape", [(5, 5, 5), (5, 6, 7), (6, 5, 5)])
def test_mv_AN_equals_dense_toeplitz(shape):
    rng = np.random.default_rng(sum(shape))
    L, M, N = shape
    G = _parity_G(L, M, N, rng)
    fN = oop.fft_operator(oop.circulant_nop(G))
    Mr = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    Mc = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    idx = np.sort(rng.choice(3 * L * M * N, 150, replace=False))
    x = rng.standard_normal(150) + 1j * rng.standard_normal(150)
    y = mvp.mv_AN(x, fN, Mr, Mc, 0.7, 'notransp', idx)
    yd = mvp.toeplitz_apply_dense(x, G, Mr, Mc, 0.7, idx)
    assert np.linalg.norm(y - yd) <= 1e-13 * np.linalg.norm(yd)


def test_transp_is_adjoint_and_linear():
    rng = np.random.default_rng(3)
    L, M, N = 6, 5, 7
    fN = rng.standard_normal((2 * L, 2 * M, 2 * N, 6)) + 1j * rng.standard_normal((2 * L, 2 * M, 2 * N, 6))
    G = _parity_G(L, M, N, rng)
    fN = oop.fft_operator(oop.circulant_nop(G))
    Mr = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    Mc = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    idx = np.arange(3 * L * M * N)
    u = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    v = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    Av = mvp.mv_AN(v, fN, Mr, Mc, 0.3, 'notransp', idx)
    AHu = mvp.mv_AN(u, fN, Mr, Mc, 0.3, 'transp', idx)
    assert abs(np.vdot(AHu, v) - np.vdot(u, Av)) <= 1e-12 * abs(np.vdot(u, Av))
    a, b = 0.3 - 2j, 1.7 + 0.1j
    lin = mvp.mv_AN(a * u + b * v, fN, Mr, Mc, 0.3, 'notransp', idx)
    assert np.linalg.norm(lin - a * mvp.mv_AN(u, fN, Mr, Mc, 0.3, 'notransp', idx) - b * Av) <= 1e-13 * np.linalg.norm(lin)


def test_mv_L_and_N_relation():
    rng = np.random.default_rng(4)
    L, M, N = 5, 5, 6
    fN = rng.standard_normal((2 * L, 2 * M, 2 * N, 6)) + 0j
    idx = np.arange(3 * L * M * N)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    assert np.allclose(mvp.mv_L(x, fN, L, M, N, 0.2, idx), mvp.mv_N(x, fN, L, M, N, idx) - 0.2 * x, atol=1e-13)
    # mv_AN with Mr = 1, Mc = -1 equals Gram x + N x
    ones = np.ones((L, M, N))
    assert np.allclose(mvp.mv_AN(x, fN, ones, -ones, 0.2, 'notransp', idx),
                       0.2 * x + mvp.mv_N(x, fN, L, M, N, idx), atol=1e-12)


def _givens_gmres(A, b, m, tol, maxit):
    """Independent textbook GMRES(m) with Givens rotations (Saad, Alg. 6.9)."""
    n = b.size
    x = np.zeros(n, complex)
    bn = np.linalg.norm(b)
    hist = [1.0]
    for _ in range(maxit):
        r = b - A(x)
        beta = np.linalg.norm(r)
        if beta / bn <= tol:
            break
        V = np.zeros((n, m + 1), complex)
        H = np.zeros((m + 1, m), complex)
        V[:, 0] = r / beta
        g = np.zeros(m + 1, complex)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m, complex)
        for k in range(m):
            w = A(V[:, k])
            for i in range(k + 1):
                H[i, k] = np.vdot(V[:, i], w)
            w = w - V[:, :k + 1] @ H[:k + 1, k]
            H[k + 1, k] = np.linalg.norm(w)
            V[:, k + 1] = w / H[k + 1, k]
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -np.conj(sn[i]) * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            a_, b_ = H[k, k], H[k + 1, k]
            t = np.hypot(abs(a_), abs(b_))
            cs[k] = abs(a_) / t
            sn[k] = (a_ / abs(a_)) * np.conj(b_) / t
            H[k, k] = cs[k] * a_ + sn[k] * b_
            H[k + 1, k] = 0
            g[k + 1] = -np.conj(sn[k]) * g[k]
            g[k] = cs[k] * g[k]
            hist.append(abs(g[k + 1]) / bn)
            if abs(g[k + 1]) < tol * bn:
                k += 1
                break
        else:
            k = m
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + V[:, :k] @ y
    return x, np.array(hist)


def test_pgmres_matches_dense_solve_and_givens():
    rng = np.random.default_rng(5)
    n = 60
    A = np.eye(n) * 4 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = lambda v: A @ v
    x, flag, relres, it, resvec = krylov.pgmres(f, b, 10, 1e-10, 50)
    xs = np.linalg.solve(A, b)
    assert flag == 0 and relres <= 1e-10
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    xg, hist = _givens_gmres(f, b, 10, 1e-10, 50)
    # identical residual history in exact arithmetic; same at the first cycle numerically
    assert np.allclose(resvec[:11], hist[:11], rtol=1e-8, atol=1e-14)
    assert resvec[0] == 1.0
    assert it[0] >= 1 and it[1] >= 1


def test_pgmres_flags_and_x0():
    rng = np.random.default_rng(6)
    n = 40
    A = np.eye(n) + 0.5 * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 0j
    f = lambda v: A @ v
    x, flag, relres, it, resvec = krylov.pgmres(f, b, 3, 1e-12, 1)
    assert flag == 1 and it[1] == 1 and len(resvec) == 4
    xs = np.linalg.solve(A, b)
    x, flag, relres, it, resvec = krylov.pgmres(f, b, 20, 1e-8, 50, x0=xs)
    assert flag == 0 and len(resvec) == 1        # converged initial guess: no iteration


def test_bicgstab_converges():
    rng = np.random.default_rng(7)
    n = 50
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, flag, relres, it, resvec = krylov.bicgstab(lambda v: A @ v, b, 1e-10, 200)
    assert flag == 0 and np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("k", [1, 2, 3, 6])
def test_bicgstab_iterates_match_scipy(k):
    """After k full iterations (tolerance never met) the oracle's iterate is
    scipy.sparse.linalg.bicgstab's, whose recurrences are van der Vorst's with
    the shadow residual r0 -- with and without a preconditioner."""
    import scipy.sparse.linalg as sla
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) * 2 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) * 0.8
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    P = np.diag(1 / np.diag(A)) + 0.01 * rng.standard_normal((n, n))
    xs, _ = sla.bicgstab(A, b, rtol=1e-30, atol=0, maxiter=k)
    x, flag, relres, it, resvec = krylov.bicgstab(lambda v: A @ v, b, 1e-30, k)
    assert np.linalg.norm(x - xs) <= 1e-13 * np.linalg.norm(xs)
    assert it == k and len(resvec) == 2 * k + 1 and flag == 1
    assert abs(resvec[-1] - np.linalg.norm(b - A @ x)) <= 1e-10 * resvec[0]
    xs, _ = sla.bicgstab(A, b, rtol=1e-30, atol=0, maxiter=k, M=sla.aslinearoperator(P))
    x, *_ = krylov.bicgstab(lambda v: A @ v, b, 1e-30, k, M=lambda v: P @ v)
    assert np.linalg.norm(x - xs) <= 1e-13 * np.linalg.norm(xs)


def test_bicgstab_half_step_and_breakdown():
    # A = c I: the first half step is exact (iter 0.5, x = b / c)
    b = np.array([1 + 1j, 2.0, -0.5j])
    x, flag, relres, it, resvec = krylov.bicgstab(lambda v: (2 - 1j) * v, b, 1e-12, 10)
    assert flag == 0 and it == 0.5 and len(resvec) == 2
    assert np.allclose(x, b / (2 - 1j), rtol=1e-15, atol=0)
    # real skew-symmetric A: <r0, A r0> = 0 breaks down before any update
    A = np.array([[0.0, 1.0], [-1.0, 0.0]])
    x, flag, relres, it, resvec = krylov.bicgstab(lambda v: A @ v, np.array([1.0, 2.0]), 1e-10, 10)
    assert flag == 4 and it == 0 and relres == 1.0 and np.all(x == 0)


def _dr_matrix(n=90, seed=1):
    """A non-normal complex matrix with four small eigenvalues (what deflation removes)."""
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    T = np.triu(0.05 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))), 1)
    ev = np.concatenate([[1e-3, 2e-3, 5e-3, 1e-2], 1 + rng.random(n - 4)]) * np.exp(0.3j * rng.standard_normal(n))
    A = Q @ (np.diag(ev) + T) @ Q.conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def test_pgmresdr_solves_and_first_cycle_is_gmres():
    A, b = _dr_matrix()
    xe = np.linalg.solve(A, b)
    Af = lambda v: A @ v
    x, flag, relres, it, rv = krylov.pgmresDR(Af, b, 20, 1e-10, 100, 5)
    assert flag == 0 and relres <= 1e-10 and np.linalg.norm(x - xe) <= 1e-8 * np.linalg.norm(xe)
    _, _, _, _, rvg = krylov.pgmres(Af, b, 20, 1e-10, 100)
    assert np.allclose(rv[:21], rvg[:21], rtol=1e-10, atol=0)      # the first cycle is GMRES(20)
    assert rv.size < rvg.size                                       # deflation pays
    P = np.diag(1 / np.diag(A))
    for kw in (dict(R1=lambda v: P @ v), dict(L1=lambda v: P @ v)):
        x, flag, relres, it, rv = krylov.pgmresDR(Af, b, 20, 1e-10, 100, 5, **kw)
        assert flag == 0 and np.linalg.norm(x - xe) <= 1e-8 * np.linalg.norm(xe)


def test_harmonic_ritz_vectors_pinned():
    """getHarmVecs1: the eigenvalues are the harmonic Ritz values of the cycle,
    the pencil (Hbar^H Hbar, H_m^H); getHarmVecs2: 1/theta for the harmonic
    Ritz values of the dense A over span([U V_p]), and A [U V_p] = [C V] H2."""
    import scipy.linalg as sl
    A, b = _dr_matrix(60, 4)
    Af = lambda v: A @ v
    m, k = 12, 4
    V, H, _, p, _ = krylov.pgmresDR_iter(Af, b.astype(complex), m, None, None, None, None, None, 0.0)
    P, vals = krylov.harm_vecs1(p, k, H)
    theta = sl.eig(H.conj().T @ H, H[:p].conj().T, right=False)
    ref = np.sort_complex(theta[np.argsort(np.abs(theta))[:k]])
    assert np.allclose(np.sort_complex(vals), ref, rtol=1e-9)
    # the first recycled pair, then one more cycle against it
    y = np.linalg.lstsq(H, V.conj().T @ b, rcond=None)[0]
    r = b - V @ (H @ y)
    U = V[:, :p] @ P
    Q, R = np.linalg.qr(H @ P)
    C = V @ Q
    U = np.linalg.solve(R.T, U.T).T
    assert np.allclose(A @ U, C, atol=1e-10)
    V, H, B, p, _ = krylov.pgmresDR_iter(Af, r, m, None, None, None, None, C, 0.0)
    D = np.diag(1 / np.linalg.norm(U, axis=0))           # pgmresDR.m:223-227
    U = U @ D
    H2 = np.block([[D, B], [np.zeros((H.shape[0], k)), H]])
    W = np.hstack([U, V[:, :p]])
    CV = np.hstack([C, V])
    assert np.allclose(A @ W, CV @ H2, atol=1e-9)
    P2, vals2 = krylov.harm_vecs2(p + k, k, H2, V, U, C)
    th = sl.eig(W.conj().T @ A.conj().T @ A @ W, W.conj().T @ A.conj().T @ W, right=False)
    ref2 = np.sort_complex(th[np.argsort(np.abs(th))[:k]])
    assert np.allclose(np.sort_complex(1 / vals2), ref2, rtol=1e-8)
