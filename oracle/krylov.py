"""Oracle: restarted GMRES exactly as MARIE's pgmres / pgmres_iter.

Restates src_numeric/src_iterative/pgmres.m:70-181 and pgmres_iter.m:22-150:
classical Gram-Schmidt in one pass (pgmres_iter.m:106-107), the small
least-squares problem solved from scratch at every inner step
(pgmres_iter.m:113-120), convergence on the LS residual against tol*||b||,
restart loop on resvec, and flag/relres from the TRUE residual
||b - A x|| / ||b|| (pgmres.m:175-180). Preconditioner handles return the
preconditioned vector directly (iterapp 'mtimes' on a function handle).
Also GMRES-DR (pgmresDR.m, getHarmVecs1/2.m) and BiCGStab (MATLAB's
built-in, not in the reference tree: its iterates are pinned by scipy's
BiCGStab, while MATLAB's exact iterate sequence and internal stagnation
tests are parity unpinned).
Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np


def pgmres_iter(A, x, r, m, L1, L2, R1, R2, tol):
    n = r.shape[0]
    V = np.zeros((n, m + 1), complex)
    V[:, 0] = r / np.linalg.norm(r)
    H = np.zeros((m + 1, m), complex)
    resvec = np.zeros(m)
    rnorm = np.linalg.norm(r)
    y = None
    res = None
    k = 0
    for k in range(1, m + 1):
        w = V[:, k - 1]
        if R1 is not None:
            w = R1(w)
        if R2 is not None:
            w = R2(w)
        w = A(w)
        if L1 is not None:
            w = L1(w)
        if L2 is not None:
            w = L2(w)
        h = V[:, :k].conj().T @ w
        H[:k, k - 1] = h
        w = w - V[:, :k] @ h
        H[k, k - 1] = np.linalg.norm(w)
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, complex)
        rhs[0] = rnorm
        Hc = H[:k + 1, :k]
        y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]
        res = rhs - Hc @ y
        resvec[k - 1] = np.linalg.norm(res)
        if resvec[k - 1] < tol:
            break
    resvec = resvec[:k]
    xupd = V[:, :k] @ y
    if R1 is not None:
        xupd = R1(xupd)
    if R2 is not None:
        xupd = R2(xupd)
    x = x + xupd
    r = V[:, :k + 1] @ res
    return x, r, k, resvec


def pgmres(A, b, restart=50, tol=1e-3, maxit=10, L1=None, L2=None, R1=None, R2=None, x0=None):
    """Returns (x, flag, relres, iter, resvec) with iter = [inner, outer] (pgmres.m:70-181)."""
    if x0 is None:
        x0 = np.zeros_like(b, dtype=complex)
    if np.count_nonzero(x0):
        r = b - A(x0)
    else:
        r = b.astype(complex)
    bp = b
    if L1 is not None:
        r = L1(r); bp = L1(bp)
    if L2 is not None:
        r = L2(r); bp = L2(bp)
    bnorm = np.linalg.norm(bp)
    resvec = [np.linalg.norm(r) / bnorm]
    x = x0.astype(complex)
    outit = 0
    p = 0
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        x, r, p, rv = pgmres_iter(A, x, r, restart, L1, L2, R1, R2, tol * bnorm)
        resvec.extend(list(rv / bnorm))
    relres = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [p, outit], np.array(resvec)


def pgmresDR_iter(A, r, m, L1, L2, R1, R2, C, tol):
    """One cycle of pgmresDR_iter.m:22-142: Arnoldi on A (right preconditioners
    R2 then R1, left L1 then L2), each new vector first orthogonalised against
    the recycled basis C (B(:,k) = C^H w, :105-109) then against V by one pass
    of classical Gram-Schmidt; the least-squares residual of [H] y = ||r|| e1
    is the stopping test. Returns V (n x k+1), H (k+1 x k), B (kc x k), k,
    resvec."""
    n = r.shape[0]
    V = np.zeros((n, m + 1), complex)
    rnorm = np.linalg.norm(r)
    V[:, 0] = r / rnorm
    kc = 0 if C is None else C.shape[1]
    B = np.zeros((kc, m), complex)
    H = np.zeros((m + 1, m), complex)
    resvec = []
    for k in range(1, m + 1):
        w = V[:, k - 1]
        if R2 is not None:
            w = R2(w)
        if R1 is not None:
            w = R1(w)
        w = A(w)
        if L1 is not None:
            w = L1(w)
        if L2 is not None:
            w = L2(w)
        if kc:
            B[:, k - 1] = C.conj().T @ w
            w = w - C @ B[:, k - 1]
        h = V[:, :k].conj().T @ w
        H[:k, k - 1] = h
        w = w - V[:, :k] @ h
        H[k, k - 1] = np.linalg.norm(w)
        V[:, k] = w / H[k, k - 1]
        rhs = np.zeros(k + 1, complex)
        rhs[0] = rnorm
        Hc = H[:k + 1, :k]
        y = np.linalg.lstsq(Hc, rhs, rcond=None)[0]
        resvec.append(np.linalg.norm(rhs - Hc @ y))
        if resvec[-1] < tol:
            return V[:, :k + 1], H[:k + 1, :k], B[:, :k], k, np.array(resvec)
    return V, H, B, m, np.array(resvec)


def harm_vecs1(m, k, H):
    """getHarmVecs1.m:14-28: eigenvectors of H_m + h_{m+1,m}^2 H_m^{-H} e_m e_m^T
    (the harmonic Ritz problem of a GMRES cycle) for the k eigenvalues of
    smallest magnitude. Returns (vectors m x k, the k eigenvalues)."""
    Hm = H[:m, :]
    Z = np.linalg.solve(Hm.conj().T, np.eye(m))
    Z[:m, :m - 1] = 0
    Mat = Hm + H[m, m - 1] ** 2 * Z
    vals, vecs = np.linalg.eig(Mat)
    order = np.argsort(np.abs(vals), kind='stable')[:k]
    return vecs[:, order], vals[order]


def harm_vecs2(m, k, H2, V, U, C):
    """getHarmVecs2.m:18-35: the pencil (H2^H [C^H U, 0; V^H U, I], H2^H H2)
    whose eigenvalues are 1/theta for the harmonic Ritz values theta of A over
    span([U V_p]); the k of largest magnitude (smallest |theta|). Returns
    (vectors m x k, their eigenvalues)."""
    import scipy.linalg as sl
    Bm = H2.conj().T @ H2
    Am = np.zeros((m + 1, m), complex)
    Am[:k, :k] = C.conj().T @ U
    Am[k:m + 1, :k] = V.conj().T @ U
    Am[k:m, k:m] = np.eye(m - k)
    Am = H2.conj().T @ Am
    vals, vecs = sl.eig(Am, Bm)
    order = np.argsort(np.abs(vals), kind='stable')[m - k:]
    return vecs[:, order], vals[order]


def pgmresDR(A, b, restart=50, tol=1e-3, maxit=10, ritz=None, L1=None, L2=None, R1=None, R2=None, x0=None):
    """GMRES with deflated restarts as pgmresDR.m:73-282 (a GCRO-DR cycle:
    the recycled pair U, C = A U with C orthonormal, harmonic Ritz vectors of
    smallest magnitude kept between cycles). Returns (x, flag, relres,
    [inner, outer], resvec), resvec relative to ||L b||.

    The first cycle is plain GMRES(restart) followed by the least-squares
    update y = H \\ (V^H r) (:175-177) and getHarmVecs1 (:180-197); later
    cycles solve the (k + p + 1) x (k + p) system H2 = [D B; 0 H] (:223-235)
    and refresh U, C with getHarmVecs2 (:238-250). When the first cycle stops
    after p steps with p - 1 < ritz, the reference keeps p - 1 vectors but
    then indexes `ritz` columns of U (:223-224), an error; here every later
    cycle uses the number of columns U has.
    """
    if ritz is None:
        ritz = min(10, round(restart / 2))
    else:
        ritz = min(ritz, restart - 1)
    x0 = np.zeros_like(b, dtype=complex) if x0 is None else np.asarray(x0, complex)
    r = b - A(x0)
    bp = b
    if L1 is not None:
        r = L1(r); bp = L1(bp)
    if L2 is not None:
        r = L2(r); bp = L2(bp)
    bnorm = np.linalg.norm(bp)
    resvec = [np.linalg.norm(r) / bnorm]
    outit = 0
    V, H, _, p, rv = pgmresDR_iter(A, r, restart, L1, L2, R1, R2, None, tol * bnorm)
    resvec.extend(list(rv / bnorm))
    y = np.linalg.lstsq(H, V.conj().T @ r, rcond=None)[0]
    x = V[:, :p] @ y
    r = r - V @ (H @ y)
    U = C = None
    if resvec[-1] > tol:
        k = p - 1 if ritz > p - 1 else ritz
        P, _ = harm_vecs1(p, k, H)
        U = V[:, :p] @ P
        Q, R = np.linalg.qr(H @ P)
        C = V @ Q
        U = np.linalg.solve(R.T, U.T).T          # U / R
    while resvec[-1] > tol and outit < maxit:
        outit += 1
        V, H, B, p, rv = pgmresDR_iter(A, r, restart, L1, L2, R1, R2, C, tol * bnorm)
        resvec.extend(list(rv / bnorm))
        k = U.shape[1]
        d = np.linalg.norm(U, axis=0)
        D = np.diag(1.0 / d)                     # pgmresDR.m:223-227
        U = U @ D
        H2 = np.block([[D, B], [np.zeros((H.shape[0], k)), H]])
        CV = np.hstack([C, V])
        y = np.linalg.lstsq(H2, CV.conj().T @ r, rcond=None)[0]
        x = x + np.hstack([U, V[:, :p]]) @ y
        r = r - CV @ (H2 @ y)
        if resvec[-1] > tol:
            P, _ = harm_vecs2(p + k, k, H2, V, U, C)
            U = np.hstack([U, V[:, :p]]) @ P
            Q, R = np.linalg.qr(H2 @ P)
            C = CV @ Q
            U = np.linalg.solve(R.T, U.T).T
    if R2 is not None:
        x = R2(x)
    if R1 is not None:
        x = R1(x)
    x = x0 + x
    relres = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    flag = 0 if relres <= tol else 1
    return x, flag, relres, [p, outit], np.array(resvec)


def bicgstab(A, b, tol=1e-6, maxit=None, M=None, x0=None):
    """Preconditioned BiCGStab (van der Vorst 1992) with MATLAB's return
    convention (x, flag, relres, iter, resvec), the solver of
    JVIE_Solver.m:463-490 (bicgstab(afun, Vrhs, tol, max_it[, pfun]); the
    preconditioner handle returns the preconditioned vector, iterapp.m:42-58).

    MATLAB's built-in bicgstab is not in the reference tree, so this follows
    its documented contract: the shadow residual is r0; resvec holds the
    absolute residual norm at every half step starting with ||r0||; the
    iteration stops when a half- or full-step residual is <= tol*||b|| and
    iter counts half steps (k - 0.5 when it stops halfway); a zero <r~, r>,
    <r~, v>, <t, t> or omega ends it as a breakdown (flag 4 unless the true
    residual meets tol); flag and relres come from the true residual.
    Its iterates are pinned by scipy.sparse.linalg.bicgstab (the same
    recurrences) at fixed iteration counts in the tests.
    """
    n = b.shape[0]
    maxit = maxit or min(n, 20)
    Mi = (lambda v: v) if M is None else M
    x = np.zeros(n, complex) if x0 is None else np.array(x0, complex)
    r = b - A(x) if np.count_nonzero(x) else b.astype(complex)
    nb = np.linalg.norm(b)
    rt = r.copy()
    resvec = [np.linalg.norm(r)]
    half = 0
    breakdown = False
    if resvec[0] > tol * nb:
        rho = np.vdot(rt, r)
        rho_old = alpha = omega = 1.0 + 0j
        v = np.zeros(n, complex)
        p = np.zeros(n, complex)
        for it in range(1, maxit + 1):
            if rho == 0:
                breakdown = True
                break
            if it == 1:
                p = r.copy()
            else:
                beta = (rho / rho_old) * (alpha / omega)
                p = r + beta * (p - omega * v)
            ph = Mi(p)
            v = A(ph)
            rtv = np.vdot(rt, v)
            if rtv == 0:
                breakdown = True
                break
            alpha = rho / rtv
            s = r - alpha * v
            x = x + alpha * ph
            resvec.append(np.linalg.norm(s))
            half += 1
            if resvec[-1] <= tol * nb:
                break
            sh = Mi(s)
            t = A(sh)
            tt = np.vdot(t, t)
            if tt == 0:
                breakdown = True
                break
            omega = np.vdot(t, s) / tt
            if omega == 0:
                breakdown = True
                break
            x = x + omega * sh
            r = s - omega * t
            resvec.append(np.linalg.norm(r))
            half += 1
            if resvec[-1] <= tol * nb:
                break
            rho_old = rho
            rho = np.vdot(rt, r)
    relres = np.linalg.norm(b - A(x)) / nb if nb > 0 else 0.0
    flag = 0 if relres <= tol else (4 if breakdown else 1)
    return x, flag, relres, half / 2, np.array(resvec)
