"""Oracle: restarted GMRES exactly as MARIE's pgmres / pgmres_iter.

Restates src_numeric/src_iterative/pgmres.m:70-181 and pgmres_iter.m:22-150:
classical Gram-Schmidt in one pass (pgmres_iter.m:106-107), the small
least-squares problem solved from scratch at every inner step
(pgmres_iter.m:113-120), convergence on the LS residual against tol*||b||,
restart loop on resvec, and flag/relres from the TRUE residual
||b - A x|| / ||b|| (pgmres.m:175-180). Preconditioner handles return the
preconditioned vector directly (iterapp 'mtimes' on a function handle).
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
