"""Oracle: adjoint gradient of a real objective w.r.t. per-voxel eps_r and sigma.

Test infrastructure only (see oracle/__init__.py). The reference has no
derivatives; this restates the adjoint of its forward JVIE solve
(JVIE_Solver.m:131-156, form II :242-244; e_r = eps_r - j sigma/(eo omega),
Scat_Solver.m:144) so the device gradient has a CPU counterpart:

  F(x, Mc) = A x - b = Gram x - Mc (N x + j omega eo Gram Einc),  Mc = (e_r - 1)/e_r
  dMc/de_r = (1 - Mc)^2
  A^H lambda = g = dJ/dconj(x)            (mv_AN 'transp', mv_AN.m:65-98)
  S_v = sum_c conj(lambda_vc) (1 - Mc_v)^2 u_vc,  u = N x + j omega eo Gram Einc
  dJ/d eps_r = 2 Re S_v,  dJ/d sigma = 2 Im S_v / (omega eo)

Pinned by central finite differences of the oracle's own forward solve
(tests/test_oracle_adjoint.py).
"""
import numpy as np

from . import fields, krylov, mvp


def forward(eps_r, sigma, Einc, dx, freq, fN, tol=1e-13, restart=50, maxit=200):
    """Form II solve for per-voxel eps_r, sigma (L x M x N, body where eps_r != 1
    or sigma != 0); returns x on the DOF index, the index, Mc and Einc on the DOFs."""
    em = fields.em_constants(freq)
    e_r = eps_r - 1j * sigma / (em['eo'] * em['omega'])
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, freq, form=2)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    x, flag, relres, it, rv = krylov.pgmres(A, b, restart, tol, maxit)
    einc_dof = np.asarray(Einc).ravel(order='F')[idx]
    return x, idx, Mc, einc_dof, Gram, em


def adjoint_state(g, fN, Mc, Gram, idx, tol=1e-13, restart=50, maxit=200):
    """lambda with A^H lambda = g."""
    Mr = np.ones_like(Mc)
    AH = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'transp', idx)
    lam, flag, relres, it, rv = krylov.pgmres(AH, g, restart, tol, maxit)
    return lam


def material_gradient(lam, x, einc_dof, fN, Mc, Gram, idx, omega_eo):
    """(dJ/deps_r, dJ/dsigma) per voxel (L x M x N arrays, 0 outside the DOFs)."""
    L, M, N = Mc.shape
    n = L * M * N
    u = mvp.mv_N(x, fN, L, M, N, idx) + 1j * omega_eo * Gram * einc_dof
    S = np.zeros(n, complex)
    vox = idx % n
    np.add.at(S, vox, np.conj(lam) * u)
    S = S * (1.0 - Mc.ravel(order='F')) ** 2
    ge = (2.0 * S.real).reshape(L, M, N, order='F')
    gs = (2.0 * S.imag / omega_eo).reshape(L, M, N, order='F')
    return ge, gs
