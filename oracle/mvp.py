"""Oracle: FFT-circulant matvecs of the N operator.

Restates src_vie/src_mvp/mv_AN.m:1-147, mv_N.m:1-95 and mv_L.m:1-98.
`idx` is 0-based into the column-major L x M x N x 3 grid (MATLAB's idx - 1).
Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np
import scipy.fft as sfft

# (output component a, input component c) -> fN component (mv_AN.m:108-123):
# the symmetric tensor [[xx,xy,xz],[xy,yy,yz],[xz,yz,zz]].
TENSOR = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 5]])


def _scatter(x, idx, L, M, N):
    J = np.zeros(3 * L * M * N, complex)
    J[idx] = x
    return J.reshape(L, M, N, 3, order='F')


def _n_apply(J, fN, workers=None):
    """crop(ifftn(fN (x) fftn(pad(J)))) for the 3 components (mv_AN.m:108-131)."""
    L, M, N, _ = J.shape
    P = fN.shape[:3]
    fJ = [sfft.fftn(J[..., c], s=P, axes=(0, 1, 2), workers=workers) for c in range(3)]
    out = np.empty((L, M, N, 3), complex)
    for a in range(3):
        acc = fN[..., TENSOR[a, 0]] * fJ[0]
        acc += fN[..., TENSOR[a, 1]] * fJ[1]
        acc += fN[..., TENSOR[a, 2]] * fJ[2]
        out[..., a] = sfft.ifftn(acc, axes=(0, 1, 2), workers=workers)[:L, :M, :N]
    return out


def mv_AN(x, fN, Mr, Mc, Gram, transp_flag, idx, workers=None):
    """JOut = A_N x (notransp) or A_N^H x (transp) over the compact DOFs idx.

    notransp: Gram*Mr*J - Mc*N(J)                 (mv_AN.m:101-131)
    transp:   Gram*conj(Mr)*J - conj(N(Mc*conj J)) (mv_AN.m:65-98)
    """
    L, M, N = Mr.shape
    J = _scatter(x, idx, L, M, N)
    Mr3 = Mr[..., None]
    Mc3 = Mc[..., None]
    if transp_flag == 'transp':
        y = _n_apply(np.conj(np.conj(Mc3) * J), fN, workers)
        out = Gram * np.conj(Mr3) * J - np.conj(y)
    elif transp_flag == 'notransp':
        y = _n_apply(J, fN, workers)
        out = Gram * Mr3 * J - Mc3 * y
    else:
        raise ValueError(transp_flag)
    return out.ravel(order='F')[idx]


def mv_N(x, fN, L, M, N, idx, workers=None):
    """JOut = N x (mv_N.m:55-95)."""
    J = _scatter(x, idx, L, M, N)
    return _n_apply(J, fN, workers).ravel(order='F')[idx]


def mv_L(x, fN, L, M, N, Gram, idx, workers=None):
    """JOut = N x - Gram x (mv_L.m:55-98)."""
    J = _scatter(x, idx, L, M, N)
    out = _n_apply(J, fN, workers) - Gram * J
    return out.ravel(order='F')[idx]


def mv_K(x, fK, L, M, N, idx, workers=None):
    """JOut = K x (mv_K.m:55-95): per component the FFT of the input, the cross
    product with the three fK components, inverse FFT and crop."""
    J = _scatter(x, idx, L, M, N)
    P = fK.shape[:3]
    F = [sfft.fftn(J[..., c], s=P, workers=workers) for c in range(3)]
    K = [fK[..., c] for c in range(3)]
    out = [K[2] * F[1] - K[1] * F[2], -K[2] * F[0] + K[0] * F[2], K[1] * F[0] - K[0] * F[1]]
    Y = np.stack([sfft.ifftn(o, workers=workers)[:L, :M, :N] for o in out], axis=-1)
    return Y.ravel(order='F')[idx]


def toeplitz_apply_dense_k(x, G, idx):
    """Brute-force K x with the Toeplitz generator G (L x M x N x 3): the
    kernel at offset d = m - n is G(|d|) with component c negated when d_c < 0
    (the reflections of circulant_kop.m), applied as out = x cross-like
    [K3 x2 - K2 x3, -K3 x1 + K1 x3, K2 x1 - K1 x2] (mv_K.m:66-80)."""
    L, M, N, _ = G.shape
    J = _scatter(x, idx, L, M, N)
    Y = np.zeros_like(J)
    for m in np.ndindex(L, M, N):
        acc = np.zeros(3, complex)
        for nn in np.ndindex(L, M, N):
            d = np.subtract(m, nn)
            K = G[abs(d[0]), abs(d[1]), abs(d[2])].copy()
            for c in range(3):
                if d[c] < 0:
                    K[c] = -K[c]
            xv = J[nn]
            acc += np.array([K[2] * xv[1] - K[1] * xv[2], -K[2] * xv[0] + K[0] * xv[2], K[1] * xv[0] - K[0] * xv[1]])
        Y[m] = acc
    return Y.ravel(order='F')[idx]


def toeplitz_apply_dense(x, G, Mr, Mc, Gram, idx):
    """Brute-force A_N x from the Toeplitz generator G (L x M x N x 6).

    y_a(m) = sum_n sum_c T_ac(m - n) J_c(n) with T(-d) obtained by the
    per-axis parity of each component (gperiodic_coeff_nop.m:6-30). Used only
    on tiny grids to pin the FFT-circulant path.
    """
    from .operators import GP_COEFF
    L, M, N, _ = G.shape
    J = _scatter(x, idx, L, M, N)
    sx, sy, sz = GP_COEFF['L'], GP_COEFF['M'], GP_COEFF['N']
    out = np.zeros((L, M, N, 3), complex)
    ii, jj, kk = np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij')
    for i in range(L):
        for j in range(M):
            for k in range(N):
                di, dj, dk = i - ii, j - jj, k - kk
                T = G[np.abs(di), np.abs(dj), np.abs(dk)]          # L,M,N,6
                sgn =(np.where((di < 0)[..., None], sx, 1.0) * np.where((dj < 0)[..., None], sy, 1.0)
                       * np.where((dk < 0)[..., None], sz, 1.0))
                T = T * sgn
                for a in range(3):
                    out[i, j, k, a] = sum(np.sum(T[..., TENSOR[a, c]] * J[..., c]) for c in range(3))
    res = Gram * Mr[..., None] * J - Mc[..., None] * out
    return res.ravel(order='F')[idx]
