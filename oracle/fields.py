"""Oracle: EM constants, plane-wave excitation and the JVIE solve driver.

Restates src_vie/src_solver/EMconstants.m:13-20,
src_vie/src_fields/src_excitations/PlaneWave_Excitation.m:39-95 and the
form-I/II set-up of src_vie/src_solver/JVIE_Solver.m:131-156, :176-299,
:306-377, :513-514. Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np

from . import krylov, mvp


def em_constants(freq):
    mu = 4 * np.pi * 1e-7
    co = 299792458.0
    eo = 1.0 / co ** 2 / mu
    omega = 2 * np.pi * freq
    lam = co / freq
    ko = 2 * np.pi / lam
    return dict(mu=mu, co=co, eo=eo, omega=omega, lam=lam, ko=ko, omega_mu=omega * mu,
                eta=3.767303134617706e+002)


def plane_wave_excitation(r, k, omega_mu, Eo):
    """Voxel-averaged plane wave E, H (L x M x N x 3) (PlaneWave_Excitation.m:39-95)."""
    L, M, N, _ = r.shape
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    dy = r[0, 1, 0, 1] - r[0, 0, 0, 1]
    dz = r[0, 0, 1, 2] - r[0, 0, 0, 2]
    k = np.asarray(k, float)
    Eo = np.asarray(Eo, complex)
    Ho = np.cross(k, Eo) / omega_mu
    fac = []
    for ax, (kk, d) in enumerate(zip(k, (dx, dy, dz))):
        if kk != 0:
            fac.append(np.exp(-1j * kk * r[..., ax]) * (np.exp(-1j * kk * d / 2) - np.exp(1j * kk * d / 2))
                       / (-1j * kk))
        else:
            fac.append(d * np.ones((L, M, N)))
    Eexc = fac[0] * fac[1] * fac[2]
    vol = dx * dy * dz
    E = np.stack([Eo[c] * Eexc for c in range(3)], axis=-1) / vol
    H = np.stack([Ho[c] * Eexc for c in range(3)], axis=-1) / vol
    return E, H


def jvie_setup(Einc, e_r, dx, freq, form=2):
    """Mr, Mc, Gram, idxS3 (0-based) and Vrhs (JVIE_Solver.m:131-156, :242-244)."""
    em = em_constants(freq)
    L, M, N = e_r.shape
    nD = L * M * N
    Mr = e_r.astype(complex).copy()
    Mc = e_r - 1.0
    Gram = dx ** 3
    idxS = np.nonzero(np.abs(Mc.ravel(order='F')) > 1e-12)[0]
    idxS3 = np.concatenate([idxS, nD + idxS, 2 * nD + idxS])
    tau = 1j * em['omega'] * em['eo'] * Mc
    tau3 = np.concatenate([tau.ravel(order='F')] * 3)
    Vrhs = Gram * tau3[idxS3] * Einc.ravel(order='F')[idxS3]
    if form != 1:
        Mrf = Mr.ravel(order='F')
        Mc = Mc / Mr
        Vrhs = Vrhs / np.concatenate([Mrf[idxS]] * 3)
        Mr = Mr / Mr
    return Mr, Mc, Gram, idxS3, Vrhs


def jvie_solve_gmres(Einc, e_r, dx, freq, fN, tol=1e-3, form=2, precond=None,
                     inner_it=50, outer_it=200, workers=None):
    """JVIE_Solver with solver 'G' (JVIE_Solver.m:309-377); returns Je and stats."""
    Mr, Mc, Gram, idx, Vrhs = jvie_setup(Einc, e_r, dx, freq, form)
    A = lambda J: mvp.mv_AN(J, fN, Mr, Mc, Gram, 'notransp', idx, workers)
    P = lambda J: mvp.mv_AN(J, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, 'notransp', idx, workers)
    if precond == 'R':
        x, flag, relres, it, resvec = krylov.pgmres(A, Vrhs, inner_it, tol, outer_it, None, None, P)
    elif precond == 'L':
        x, flag, relres, it, resvec = krylov.pgmres(A, Vrhs, inner_it, tol, outer_it, P)
    else:
        x, flag, relres, it, resvec = krylov.pgmres(A, Vrhs, inner_it, tol, outer_it)
    L, M, N = e_r.shape
    Je = np.zeros(3 * L * M * N, complex)
    Je[idx] = x
    return Je.reshape(L, M, N, 3, order='F'), flag, relres, it, resvec
