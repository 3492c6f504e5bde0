"""MATLAB-signature mirror of MARIE's JVIE hot path on the HIP library.

Same names, argument meaning and return values as the reference functions
(indices are MATLAB's 1-based `idx`):

  mv_AN(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag)    src_vie/src_mvp/mv_AN.m:1
  mv_N(JIn0, fN, r, idx, GPU_flag)                              src_vie/src_mvp/mv_N.m:1
  mv_L(JIn0, fN, r, Gram, idx, GPU_flag)                        src_vie/src_mvp/mv_L.m:1
  getOPERATORS(r, f, op, type, singular)                        src_vie/src_operators/getOPERATORS.m:1
  PlaneWave_Excitation(r, k, omega_mu, Eo)                      .../PlaneWave_Excitation.m:1
  JVIE_Solver(Einc, e_r, r, freq, fN, tol, form, solver, ...)   src_vie/src_solver/JVIE_Solver.m:1
  EMconstants(freq)                                             src_vie/src_solver/EMconstants.m
  E_field_Nop(J, fN, Gram, freq, Einc)                          src_vie/src_fields/src_op/E_field_Nop.m:1
  mv_K(JIn0, fK, r, idx, GPU_flag)                              src_vie/src_mvp/mv_K.m:1
  H_field_Kop(J, fK, Gram, Hinc)                                src_vie/src_fields/src_op/H_field_Kop.m:1
  Scat_Solver(Einc, Hinc, freq, r, epsilon_r, sigma_e, rho, ...) src_vie/src_solver/Scat_Solver.m:1
  local_SAR / absorbed_power                                    src_vie/src_solver/Scat_Solver.m:179-228
  fVIEsolver_Assembly(RHBM, freq, gpu_flag)                     src_vie/src_solver/fVIEsolver_Assembly.m:1
  jvie2solver(Vrhs, fVIE, fVIEGPU, tol, Vini)                   src_vie/src_solver/jvie2solver.m:1

Every product runs in libmarie_vie.so on the GPU. `fN` may be the host array
returned by getOPERATORS or an already-resident `lib.Operator`; device
operators are cached per fN array so a solver loop does not re-upload it
(the reference keeps fN_gpu in a closure, JVIE_Solver.m:271-276).
GPU_flag = 0 is rejected: this package has no CPU path.
"""
import sys
import time

import numpy as np

from . import lib

C0 = 299792458.0


def EMconstants(freq):
    """mu, co, eo, omega, lambda, ko, omega_mu, eta (EMconstants.m:13-20)."""
    mu = 4 * np.pi * 1e-7
    co = C0
    eo = 1.0 / co ** 2 / mu
    omega = 2 * np.pi * freq
    lam = co / freq
    return dict(mu=mu, co=co, eo=eo, omega=omega, lam=lam, ko=2 * np.pi / lam, omega_mu=omega * mu,
                eta=3.767303134617706e+002)


_cache = {}


def _operator(fN, shape, precisions=lib.PREC_F64):
    if isinstance(fN, lib.Operator):
        return fN
    key = (id(fN), shape, precisions)
    ent = _cache.get(key)
    if ent is not None and ent[0] is fN:
        return ent[1]
    L, M, N = shape
    op = lib.Operator(L, M, N, fN.shape[:3], precisions=precisions)
    op.set_fN(fN)
    _cache.clear()
    _cache[key] = (fN, op)
    return op


def _require_gpu(GPU_flag):
    if not GPU_flag:
        raise ValueError('GPU_flag = 0: this package computes on the GPU only')


def mv_AN(JIn0, fN, Mr, Mc, Gram, transp_flag, idx, GPU_flag=1):
    _require_gpu(GPU_flag)
    Mr = np.asarray(Mr)
    op = _operator(fN, Mr.shape)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx, one_based=True)
    if transp_flag not in ('transp', 'notransp'):
        raise ValueError(transp_flag)
    return op.apply(np.asarray(JIn0).ravel(), lib.OP_AN, transp_flag == 'transp')


def mv_N(JIn0, fN, r, idx, GPU_flag=1):
    _require_gpu(GPU_flag)
    shape = np.shape(r)[:3]
    op = _operator(fN, shape)
    op.set_index(idx, one_based=True)
    return op.apply(np.asarray(JIn0).ravel(), lib.OP_N)


def mv_L(JIn0, fN, r, Gram, idx, GPU_flag=1):
    _require_gpu(GPU_flag)
    shape = np.shape(r)[:3]
    op = _operator(fN, shape)
    L, M, N = shape
    op.set_material(np.ones(shape), np.zeros(shape), Gram)
    op.set_index(idx, one_based=True)
    return op.apply(np.asarray(JIn0).ravel(), lib.OP_L)


_kcache = {}


def _k_operator(fK, shape):
    if isinstance(fK, lib.Operator):
        return fK
    key = (id(fK), shape)
    ent = _kcache.get(key)
    if ent is not None and ent[0] is fK:
        return ent[1]
    L, M, N = shape
    op = lib.Operator(L, M, N, fK.shape[:3])
    op.set_fK(fK)
    _kcache.clear()
    _kcache[key] = (fK, op)
    return op


def mv_K(JIn0, fK, r, idx, GPU_flag=1):
    """JOut = K JIn0 on the DOFs idx (1-based) (mv_K.m:1)."""
    _require_gpu(GPU_flag)
    shape = np.shape(r)[:3]
    op = _k_operator(fK, shape)
    op.set_index(idx, one_based=True)
    return op.apply(np.asarray(JIn0).ravel(), lib.OP_K)


def H_field_Kop(J, fK, Gram, Hinc=None, GPU_flag=1):
    """Total field Hinc + K J / Gram over every voxel (H_field_Kop.m:49-76)."""
    _require_gpu(GPU_flag)
    J = np.asarray(J, complex)
    if Hinc is None:
        Hinc = np.zeros_like(J)
    if not np.any(J):
        return np.asarray(Hinc, complex)
    L, M, N, _ = J.shape
    op = _k_operator(fK, (L, M, N))
    op.set_index(np.arange(1, 3 * L * M * N + 1), one_based=True)
    H = op.apply(J.ravel(order='F'), lib.OP_K).reshape(L, M, N, 3, order='F')
    return Hinc + H / Gram


def B1_plus(B):
    """B1+ = (Bx + j By) / 2, the co-rotating field component in MARIE's
    exp(+j omega t) convention (the reference defines no B1+; this is the
    usual MR definition, stated here)."""
    B = np.asarray(B)
    return (B[..., 0] + 1j * B[..., 1]) / 2


def getOPERATORS(r, f, op='N', type='FFTofC', singular='DIRECTFN', return_operator=False,
                 precisions=lib.PREC_F64, circulant_shape=None):
    """fN (2L x 2M x 2N x 6, complex128) assembled on the device.

    circulant_shape=(Px, Py, Pz), each side >= 2L-1, embeds the same Toeplitz
    operator in another circulant (e.g. 2*3*11 = 66 instead of 2*43 = 86 for
    L = 43); the matvec result is the same. By default each side is 2L unless
    2L has a prime factor above 13 (lib.circulant_side: e.g. 210 for a
    102-voxel side, where 204 = 4*3*17).

    `singular` selects the self-terms as getOPERATORS.m:36-42 does:
    'DIRECTFN' (the reference's default) evaluates the reference's DIRECTFN
    integrals on the device with Np_1D = 20 (assembly_nop.m:16), so the
    operator is the reference's default one; DIRECTFN is not converged at
    that order (its I_ST, I_EAc, I_EAo differ from the exact integrals by
    1e-5..1e-4). 'DEMCEM' uses the library's machine-precision quadrature,
    which equals DEMCEM's values to 1e-10. A length-5 complex vector supplies
    the five self-terms verbatim. With return_operator=True the
    device-resident lib.Operator is returned instead of the host array.

    type 'T' returns the Toeplitz generator (L x M x N x 6 for N, x 3 for K)
    and 'C' its circulant embedding (Px x Py x Pz x c) instead of the FFT of
    the circulant (getOPERATORS.m:88-95, :108-114); both are assembled on the
    device by the same kernels and copied back.
    """
    if op not in ('N', 'K'):
        raise NotImplementedError("operator '%s' is not on this package's path" % op)
    if type not in ('FFTofC', 'T', 'C'):
        raise ValueError("type must be 'FFTofC', 'T' or 'C' (getOPERATORS.m:88-95)")
    if type != 'FFTofC' and return_operator:
        raise ValueError("return_operator needs type 'FFTofC' (the product uses the FFT of the circulant)")
    r = np.asarray(r)
    L, M, N = r.shape[:3]
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    ko = 2 * np.pi / C0 * f                    # getOPERATORS.m:48
    o = lib.Operator(L, M, N, circulant_shape, precisions=precisions)
    if op == 'K':                              # getOPERATORS.m:101-115
        if type != 'FFTofC':
            return o.assemble_type(dx, ko, lib.OP_K, type)
        o.assemble_K(dx, ko)
        return o if return_operator else o.get_fK()
    if isinstance(singular, str):
        if singular == 'DIRECTFN':
            sv = lib.singular_terms_directfn(dx, ko, 20)
        elif singular == 'DEMCEM':
            sv = None
        else:
            raise ValueError("singular must be 'DIRECTFN', 'DEMCEM' or five complex values")
    else:
        sv = np.asarray(singular, np.complex128)
    if type != 'FFTofC':
        return o.assemble_type(dx, ko, lib.OP_N, type, sv)
    o.assemble_N(dx, ko, sv)
    if return_operator:
        return o
    return o.get_fN()


def E_field_Nop(J, fN, Gram, freq, Einc=None, GPU_flag=1):
    """Total field Einc + (N J - Gram J)/(Gram j omega eo) (E_field_Nop.m:49-84).

    J is L x M x N x 3 (the currents JVIE_Solver returns); the product is
    mv_L over every voxel (idx = 1:3LMN, E_field_Nop.m:75-76).
    """
    _require_gpu(GPU_flag)
    J = np.asarray(J, complex)
    if Einc is None:
        Einc = np.zeros_like(J)
    if not np.any(J):
        return np.asarray(Einc, complex)
    em = EMconstants(freq)
    L, M, N, _ = J.shape
    op = _operator(fN, (L, M, N))
    op.set_material(np.ones((L, M, N)), np.zeros((L, M, N)), Gram)
    op.set_index(np.arange(1, 3 * L * M * N + 1), one_based=True)
    E = op.apply(J.ravel(order='F'), lib.OP_L).reshape(L, M, N, 3, order='F')
    return Einc + E / (Gram * 1j * em['omega'] * em['eo'])


def local_SAR(J, sigma_e, rho):
    """|J|^2 / (2 sigma rho) per voxel, 0 where 2 sigma rho <= 1e-10 (Scat_Solver.m:179-186)."""
    denom = 2 * np.asarray(sigma_e, float) * np.asarray(rho, float)
    S = np.sum(np.abs(J) ** 2, axis=3)
    out = np.zeros_like(S)
    ok = np.abs(denom) > 1e-10
    out[ok] = S[ok] / denom[ok]
    return out


def absorbed_power(E, J, Gram):
    """Pabs = 0.5 Re sum(E . conj J) Gram (Scat_Solver.m:226-228)."""
    return 0.5 * np.real(np.sum(E * np.conj(J))) * Gram


def Scat_Solver(Einc, Hinc, freq, r, epsilon_r, sigma_e, rho, tol=1e-3, fN=None, fK=None, form=2, solver='M',
                precond=0, max_it=1000, inner_it=50, outer_it=200, ritz=5, gpu_flag=1, logfile=None,
                precision='fp64'):
    """[Jout, Sout, Eout, Bout, Gsar, Pabs, fN, fK] of Scat_Solver.m:1 on the device.

    e_r = eps_r - j sigma/(eo omega) (:144); solver 'M' (the default, :94-99)
    reaches JVIE_Solver as BiCGStab; operators assembled when not given
    (:157-162); currents by JVIE_Solver (:169); local SAR |J|^2/(2 sigma rho)
    (:179-186); E = E_field_Nop (:204); B = mu0 H_field_Kop (:215-216);
    Pabs = 0.5 Re sum(E . conj J) Gram and Gsar = sum(Sout) Gram (:226-229).
    """
    _require_gpu(gpu_flag)
    em = EMconstants(freq)
    e_r = np.asarray(epsilon_r) - 1j * np.asarray(sigma_e) / (em['eo'] * em['omega'])
    r = np.asarray(r)
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    Gram = dx ** 3
    if fN is None:
        fN = getOPERATORS(r, freq, 'N', return_operator=True,
                          precisions=lib.PREC_F64 | (lib.PREC_F32 if precision == 'fp32-refined' else 0))
    if fK is None:
        fK = getOPERATORS(r, freq, 'K', return_operator=True)
    Jout, *_ = JVIE_Solver(Einc, e_r, r, freq, fN, tol, form, solver, precond, max_it, inner_it, outer_it, ritz,
                           gpu_flag, logfile, precision)
    Sout = local_SAR(Jout, sigma_e, rho)
    Eout = E_field_Nop(Jout, fN, Gram, freq, Einc)
    Bout = em['mu'] * H_field_Kop(Jout, fK, Gram, Hinc)
    Pabs = absorbed_power(Eout, Jout, Gram)
    Gsar = float(np.sum(Sout) * Gram)
    return Jout, Sout, Eout, Bout, Gsar, Pabs, fN, fK


def PlaneWave_Excitation(r, k, omega_mu, Eo):
    """Voxel-averaged plane wave (PlaneWave_Excitation.m:39-95)."""
    r = np.asarray(r)
    L, M, N, _ = r.shape
    d = (r[1, 0, 0, 0] - r[0, 0, 0, 0], r[0, 1, 0, 1] - r[0, 0, 0, 1], r[0, 0, 1, 2] - r[0, 0, 0, 2])
    k = np.asarray(k, float)
    Eo = np.asarray(Eo, complex)
    Ho = np.cross(k, Eo) / omega_mu
    E = np.ones((L, M, N), complex)
    for a in range(3):
        if k[a] != 0:
            E = E * (np.exp(-1j * k[a] * r[..., a]) * (np.exp(-1j * k[a] * d[a] / 2) - np.exp(1j * k[a] * d[a] / 2))
                     / (-1j * k[a]))
        else:
            E = E * d[a]
    vol = d[0] * d[1] * d[2]
    Ef = np.stack([Eo[c] * E for c in range(3)], axis=-1) / vol
    Hf = np.stack([Ho[c] * E for c in range(3)], axis=-1) / vol
    return Ef, Hf


def jvie_system(Einc, e_r, r, freq, form=2):
    """Mr, Mc, Gram, idxS3 (1-based) and Vrhs of JVIE_Solver.m:131-156, :242-244."""
    em = EMconstants(freq)
    r = np.asarray(r)
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    L, M, N = e_r.shape
    nD = L * M * N
    Mr = np.asarray(e_r, complex).copy()
    Mc = Mr - 1.0
    Gram = dx ** 3
    idxS = np.nonzero(np.abs(Mc.ravel(order='F')) > 1e-12)[0]
    idxS3 = np.concatenate([idxS, nD + idxS, 2 * nD + idxS])
    tau = 1j * em['omega'] * em['eo'] * Mc
    Vrhs = Gram * np.tile(tau.ravel(order='F'), 3)[idxS3] * np.asarray(Einc).ravel(order='F')[idxS3]
    if form != 1:
        Mrf = Mr.ravel(order='F')
        Mc = Mc / Mr
        Vrhs = Vrhs / np.tile(Mrf[idxS], 3)
        Mr = Mr / Mr
    return Mr, Mc, Gram, idxS3 + 1, Vrhs


def JVIE_Solver(Einc, e_r, r, freq, fN, tol=1e-3, form=2, solver='B', precond=0, max_it=1000,
                inner_it=50, outer_it=200, ritz=5, gpu_flag=1, logfile=None, precision='fp64',
                inner_tol=1e-4):
    """[Je, flag, relres, iter, resvec, solvert] of JVIE_Solver.m:1 on the device.

    solver 'G': restarted GMRES (pgmres semantics), precond 'R' / 'L' with the
    preconditioner mv_AN(., 1./Mr, 1./Mr - 1) (JVIE_Solver.m:316-372);
    precision='fp32-refined' runs complex64 matvecs inside FP64 GMRES with
    FP64 iterative refinement (no preconditioner); 'D': GMRES with deflated
    restarts keeping `ritz` harmonic Ritz vectors (JVIE_Solver.m:378-450).
    Any other solver letter runs BiCGStab (JVIE_Solver.m:451-492, the default), preconditioned
    when `precond` is 'L' or 'R'; iter is then in half steps / 2 and resvec absolute.
    Forms I and II (JVIE_Solver.m:176-300). precision='fp32-refined-c64'
    stores the refinement cycles' Krylov basis in complex64 as well (every
    dot product, norm and Givens rotation still in FP64);
    precision='fp32-refined-blocks' grows that complex64 basis by locally
    orthogonalised blocks of 8 products with one block Gram-Schmidt pass per
    block (GMRES mode 3; inner_it is the cycle length, up to 503; 128 is the
    benchmark's choice).

    logfile: a path the run summary is appended to (JVIE_Solver.m:115-122,
    :159-166 and the solver line of :335, :355, :375, ...); None or a path
    that cannot be opened writes it to standard output, as fid = 1 does.
    """
    t_start = time.perf_counter()
    _require_gpu(gpu_flag)
    if tol < 1e-15 or tol >= 1:          # JVIE_Solver.m:78-80
        tol = 1e-3
    if solver not in ('G', 'D'):         # JVIE_Solver.m:85-91: anything else is BiCGStab
        solver = 'B'
    use_pc = precond in ('L', 'R')       # JVIE_Solver.m:92-98
    if use_pc and precision != 'fp64':
        raise ValueError('the preconditioned solvers run in FP64')
    modes = {'fp64': 0, 'fp32-refined': 1, 'fp32-refined-c64': 2, 'fp32-refined-blocks': 3}
    if precision not in modes:
        raise ValueError("precision must be 'fp64', 'fp32-refined', 'fp32-refined-c64' or 'fp32-refined-blocks'")
    Mr, Mc, Gram, idx, Vrhs = jvie_system(Einc, e_r, r, freq, form)
    L, M, N = e_r.shape
    prec = lib.PREC_F64 | (lib.PREC_F32 if precision != 'fp64' else 0)
    log = _open_log(logfile)
    dxv = float(np.asarray(r)[1, 0, 0, 0] - np.asarray(r)[0, 0, 0, 0])
    log.write('\n ----------------------------------------------------------')
    log.write('\n Domain:                %dx%dx%d voxels' % (L, M, N))
    log.write('\n Resolution:            %.2fmm' % (dxv * 1000))
    log.write('\n # DOFS:                %d' % (3 * L * M * N))
    log.write('\n # DOFS in Scatterer:   %d' % idx.size)
    log.write('\n Operating Frequency:   %.2f MHz' % (freq / 1e6))
    log.write('\n\n ----------------------------------------------------------')
    if isinstance(fN, lib.Operator):
        op = fN
    else:
        op = lib.Operator(L, M, N, np.shape(fN)[:3], precisions=prec)
        op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx, one_based=True)
    t0 = time.perf_counter()
    side = lib.PRECOND_NONE if not use_pc else (lib.PRECOND_RIGHT if precond == 'R' else lib.PRECOND_LEFT)
    if solver == 'D':                    # JVIE_Solver.m:378-450
        if precision != 'fp64':
            raise ValueError('GMRES-DR runs in FP64')
        x, flag, relres, it, resvec = op.gmres_dr(Vrhs, None, inner_it, tol, outer_it, ritz, side)
    elif solver == 'G':
        if use_pc:
            x, flag, relres, it, resvec = op.gmres_precond(Vrhs, None, inner_it, tol, outer_it, side=side)
        else:
            x, flag, relres, it, resvec = op.gmres(Vrhs, None, inner_it, tol, outer_it,
                                                   mode=modes[precision],
                                                   inner_tol=inner_tol)
    else:
        if precision != 'fp64':
            raise ValueError('BiCGStab runs in FP64')
        x, flag, relres, it, resvec = op.bicgstab(Vrhs, None, tol, max_it, precond=use_pc)
    solvert = time.perf_counter() - t0
    # the solver line of JVIE_Solver.m:328-495 and the closing line of :517
    pc = {'R': 'w/ Right preconditioner', 'L': 'w/ Left preconditioner'}.get(precond, 'wo/ preconditioner')
    if solver == 'G':
        head = 'GMRES solver (%d restart), %s' % (inner_it, pc)
    elif solver == 'D':
        head = 'GMRES solver (%d restart, deflated %d), %s \n' % (inner_it, ritz, pc)
    else:
        head = 'BICGSTAB solver, %s' % ('w/ preconditioner' if use_pc else 'wo/ preconditioner')
    if precision != 'fp64':
        head = head.replace('),', '), %s,' % precision, 1)
    log.write('\n GPU ' + head)
    log.write(' Time  = %.2f [sec] for %d iterations, flag %d, residue %g, relative residue %g \n'
              % (solvert, len(resvec), flag, resvec[-1] if len(resvec) else 0.0, relres))
    log.write(' VIE Solve finished. Elapsed Time  = %.2f [sec] \n' % (time.perf_counter() - t_start))
    if log is not sys.stdout:
        log.close()
    Je = np.zeros(3 * L * M * N, complex)
    Je[idx - 1] = x
    return Je.reshape(L, M, N, 3, order='F'), flag, relres, it, resvec, solvert


def _open_log(logfile):
    """fid of JVIE_Solver.m:115-122: the file opened for appending, or standard output."""
    if logfile is None:
        return sys.stdout
    try:
        return open(logfile, 'a')
    except OSError:
        return sys.stdout


def jvie2solver(Vrhs, op, tol, Vini=None, max_it=1000):
    """Vsol of jvie2solver.m:1: BiCGStab (max_it 1000, initial guess Vini) on
    the form II operator held by `op`; an all-zero right-hand side returns
    itself (jvie2solver.m:30-33)."""
    Vrhs = np.asarray(Vrhs, np.complex128)
    if np.count_nonzero(Vrhs) == 0:
        return Vrhs.copy()
    x, flag, relres, it, resvec = op.bicgstab(Vrhs, Vini, tol, max_it)
    return x


def fVIEsolver_Assembly(RHBM, freq, gpu_flag=1, block_restart=20, block_maxit=200):
    """[fVIEsolver, Gram, tau3, Scoord, dev] of fVIEsolver_Assembly.m:1.

    RHBM: dict with 'r' (L x M x N x 3), 'epsilon_r', 'sigma_e' and optionally
    'fN' / 'freqfN'. The body's bounding box plus one voxel is cut out
    (fVIEsolver_Assembly.m:82-106) and its operator assembled on the device
    when the box holds < 70 % of the grid, no fN is given or it was made at
    another frequency (:110-124); form II material (:130-135); tau3 = j omega
    eo Mc on the non-air voxels (:137-138). Scoord: the non-air voxel centres
    (:71-77); they keep the column-major order of the cut-out box.

    fVIEsolver(Vrhs, tol, Vini) solves A x = Vrhs on those voxels' DOFs: a
    vector runs jvie2solver's BiCGStab; a matrix (DOFs x ports, the per-port
    loop of VSIE_Solver.m:387-407 in one call) runs block GMRES with all
    ports in each product. `dev` is the device-resident lib.Operator.
    """
    _require_gpu(gpu_flag)
    em = EMconstants(freq)
    r = np.asarray(RHBM['r'])
    e_r = np.asarray(RHBM['epsilon_r']) - 1j * np.asarray(RHBM['sigma_e']) / (em['eo'] * em['omega'])
    dx = r[1, 0, 0, 0] - r[0, 0, 0, 0]
    Gram = dx ** 3
    L, M, N = e_r.shape
    ND = L * M * N
    idxS = np.nonzero(np.abs(e_r.ravel(order='F') - 1) > 1e-12)[0]
    Scoord = np.stack([r[..., a].ravel(order='F')[idxS] for a in range(3)], axis=1)
    lo = [r[..., a].ravel(order='F')[idxS].min() - dx for a in range(3)]
    hi = [r[..., a].ravel(order='F')[idxS].max() + dx for a in range(3)]
    idxX = np.nonzero((r[:, 0, 0, 0] >= lo[0]) & (r[:, 0, 0, 0] <= hi[0]))[0]
    idxY = np.nonzero((r[0, :, 0, 1] >= lo[1]) & (r[0, :, 0, 1] <= hi[1]))[0]
    idxZ = np.nonzero((r[0, 0, :, 2] >= lo[2]) & (r[0, 0, :, 2] <= hi[2]))[0]
    e_rnew = e_r[np.ix_(idxX, idxY, idxZ)]
    fN = RHBM.get('fN')
    if e_rnew.size < 0.7 * ND or fN is None or RHBM.get('freqfN') != freq:
        rnew = r[np.ix_(idxX, idxY, idxZ)]
        op = getOPERATORS(rnew, freq, 'N', return_operator=True)
        e_r = e_rnew
        L, M, N = e_r.shape
        ND = L * M * N
        idxS = np.nonzero(np.abs(e_r.ravel(order='F') - 1) > 1e-12)[0]
    else:
        op = fN if isinstance(fN, lib.Operator) else _operator(fN, (L, M, N))
    idxS3 = np.concatenate([idxS, ND + idxS, 2 * ND + idxS])
    Mr = e_r / e_r
    Mc = (e_r - 1.0) / e_r
    tau = 1j * em['omega'] * em['eo'] * Mc
    tau3 = np.tile(tau.ravel(order='F')[idxS], 3)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idxS3)

    def fVIEsolver(Vrhs, tol, Vini=None):
        V = np.asarray(Vrhs, np.complex128)
        if V.ndim == 1 or V.shape[1] == 1:
            x = jvie2solver(V.ravel(), op, tol, None if Vini is None else np.asarray(Vini).ravel())
            return x.reshape(V.shape)
        X = np.zeros_like(V)
        live = np.nonzero(np.any(V != 0, axis=0))[0]        # zero ports return zero (jvie2solver.m:30-33)
        if live.size:
            X0 = None if Vini is None else np.asarray(Vini, np.complex128)[:, live]
            for c0 in range(0, live.size, 64):
                cols = live[c0:c0 + 64]
                Xc, flags, relres, it, rv = op.block_gmres(V[:, cols], None if X0 is None else X0[:, c0:c0 + 64],
                                                           block_restart, tol, block_maxit)
                X[:, cols] = Xc
        return X

    return fVIEsolver, Gram, tau3, Scoord, op
