"""Oracle: N-operator (Green-tensor) assembly, circulant embedding and its FFT.

Restates src_vie/src_operators/{getOPERATORS,assembly_nop,circulant_nop,
gperiodic_coeff_nop,fft_operator,vector_projection,cube_faces,nearby_correct}.m,
cubatures/{gauss_2d,cubature_nop_sym_near}.m and mexcubatures/cubature_nop.cpp;
for the K operator assembly_kop.m, circulant_kop.m, gperiodic_coeff_kop.m,
cubatures/{cubature_kop_near_sym,nearby_correct_kop}.m, mexcubatures/cubature_kop.cpp.
Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np
import scipy.fft as sfft

from . import singular

C0 = 299792458.0

# Quadrature orders and the medium-distance region (assembly_nop.m:16-23).
NP_DEMCEM = 20
NP_1D_NEAR = 20
NP_1D_MEDIUM = 10
NP_1D_FAR = 4
N_MEDIUM = 5

# The 21 unique (observation face, source face) pairs, 0-based
# (cubature_nop.cpp:27-30; cubature_nop_sym_near.m:17-18).
INDEX_OBS = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5])
INDEX_SRC = np.array([0, 1, 2, 3, 4, 5, 0, 2, 3, 4, 5, 2, 3, 4, 5, 2, 4, 5, 4, 5, 4])
# Entries filled by translation symmetry after the 21 unique ones
# (cubature_nop.cpp:96-114): (dst_row, dst_col) <- (src_row, src_col).
SYM_FILL = [((1, 1), (0, 0)), ((2, 0), (1, 3)), ((2, 1), (0, 3)),
            ((3, 0), (1, 2)), ((3, 1), (0, 2)), ((3, 3), (2, 2)),
            ((4, 0), (1, 5)), ((4, 1), (0, 5)), ((4, 2), (3, 5)), ((4, 3), (2, 5)),
            ((5, 0), (1, 4)), ((5, 1), (0, 4)), ((5, 2), (3, 4)), ((5, 3), (2, 4)),
            ((5, 5), (4, 4))]

# Sign of each tensor component (xx,xy,xz,yy,yz,zz) in the reflected blocks of
# the circulant (gperiodic_coeff_nop.m:6-30).
GP_COEFF = {
    'L':  np.array([+1, -1, -1, +1, +1, +1], float),
    'M':  np.array([+1, -1, +1, +1, -1, +1], float),
    'N':  np.array([+1, +1, -1, +1, -1, +1], float),
    'LM': np.array([+1, +1, -1, +1, -1, +1], float),
    'LN': np.array([+1, -1, +1, +1, -1, +1], float),
    'MN': np.array([+1, -1, -1, +1, +1, +1], float),
    'LMN': np.array([+1, +1, +1, +1, +1, +1], float),
}


def gauss_1d(order):
    """Gauss-Legendre rule on [-1,1], ascending nodes (gauss_1d.m:66-1388).

    The reference stores 30-digit tables (correctly rounded doubles once
    parsed). Here numpy's leggauss nodes are polished by Newton steps on the
    Legendre recurrence in extended precision and the weights recomputed,
    which reproduces the tables to the last bit (tests compare against the
    table values kept in tests/golden/ref_cubature.npz).
    Returns (weights, nodes) in the reference's output order.
    """
    x0, _ = np.polynomial.legendre.leggauss(order)
    z = x0.astype(np.longdouble)
    n = order
    for _ in range(4):
        p1 = np.ones_like(z)
        p2 = np.zeros_like(z)
        for j in range(1, n + 1):
            p3 = p2
            p2 = p1
            p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
        dp = n * (z * p1 - p2) / (z * z - 1)
        z = z - p1 / dp
    p1 = np.ones_like(z)
    p2 = np.zeros_like(z)
    for j in range(1, n + 1):
        p3 = p2
        p2 = p1
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
    dp = n * (z * p1 - p2) / (z * z - 1)
    w = 2 / ((1 - z * z) * dp * dp)
    return w.astype(np.float64), z.astype(np.float64)


def gauss_2d(order):
    """Tensor rule mapped to [0,1]^2, u outer / v inner, weights/4 (gauss_2d.m:1-19)."""
    w1, x1 = gauss_1d(order)
    w = np.outer(w1, w1).ravel() / 4.0
    u = np.repeat((x1 + 1.0) / 2.0, order)
    v = np.tile((x1 + 1.0) / 2.0, order)
    return order * order, w, u, v


def cube_faces(dx, dy=None, dz=None):
    """R_faces[coord, vertex, face] of the unit cell (cube_faces.m:1-39).

    Faces: -x, +x, -y, +y, -z, +z.
    """
    dy = dx if dy is None else dy
    dz = dx if dz is None else dz
    R = np.zeros((3, 4, 6))
    R[:, :, 0] = np.array([[0, 0, 0], [0, dy, 0], [0, dy, dz], [0, 0, dz]]).T
    R[:, :, 1] = np.array([[dx, 0, 0], [dx, dy, 0], [dx, dy, dz], [dx, 0, dz]]).T
    R[:, :, 2] = np.array([[0, 0, 0], [dx, 0, 0], [dx, 0, dz], [0, 0, dz]]).T
    R[:, :, 3] = np.array([[0, dy, 0], [dx, dy, 0], [dx, dy, dz], [0, dy, dz]]).T
    R[:, :, 4] = np.array([[0, 0, 0], [dx, 0, 0], [dx, dy, 0], [0, dy, 0]]).T
    R[:, :, 5] = np.array([[0, 0, dz], [dx, 0, dz], [dx, dy, dz], [0, dy, dz]]).T
    return R


def vector_projection():
    """P[k,l,c] = (n_k x e_a).(n_l x e_b), c over (xx,xy,xz,yy,yz,zz) (vector_projection.m:3-25)."""
    e = np.eye(3)
    KL = np.array([-e[0], e[0], -e[1], e[1], -e[2], e[2]])
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    P = np.zeros((6, 6, 6))
    for k in range(6):
        for l in range(6):
            for c, (a, b) in enumerate(pairs):
                P[k, l, c] = np.dot(np.cross(KL[k], e[a]), np.cross(KL[l], e[b]))
    return P


def cubature_nop_batch(order, r_m, R_faces, ko, dx, chunk=4096):
    """6x6 face-pair integrals I[p,q] = dx^4 sum w_i w_j exp(-jkR)/R for many offsets.

    Restates cubature_nop.cpp:8-125 (and cubature_nop_sym_near.m:1-88, which is
    the same formula) vectorised over the observation-cell offsets r_m (K x 3);
    the source cell sits at r_n = 0 (assembly_nop.m:29).
    """
    r_m = np.atleast_2d(np.asarray(r_m, float))
    K = r_m.shape[0]
    Np, w, u, v = gauss_2d(order)
    W = np.outer(w, w)
    out = np.zeros((K, 6, 6), complex)
    with np.errstate(divide='ignore', invalid='ignore'):
        for s in range(0, K, chunk):
            rm = r_m[s:s + chunk]
            for io, iq in zip(INDEX_OBS, INDEX_SRC):
                p1 = R_faces[:, 0, io]; ap = R_faces[:, 1, io] - p1; bp = R_faces[:, 3, io] - p1
                q1 = R_faces[:, 0, iq]; aq = R_faces[:, 1, iq] - q1; bq = R_faces[:, 3, iq] - q1
                rp = p1[None, :] + np.outer(u, ap) + np.outer(v, bp)        # Np x 3
                rq = q1[None, :] + np.outer(u, aq) + np.outer(v, bq)        # Np x 3
                d = rq[None, None, :, :] - rp[None, :, None, :] - rm[:, None, None, :]
                R = np.sqrt(np.sum(d * d, axis=-1))                          # k x Np x Np
                G = np.exp(-1j * ko * R) / R
                out[s:s + chunk, io, iq] = np.einsum('kpq,pq->k', G, W)
    for (dr, dc), (sr, sc) in SYM_FILL:
        out[:, dr, dc] = out[:, sr, sc]
    return out * dx ** 4


def nearby_tables(ST, EAc, EAo, VAc, VAo):
    """C and P tables of nearby_correct.m:5-133 keyed by 0-based offset (mx,my,mz)."""
    z = 0.0
    T = {}
    T[(0, 0, 0)] = (
        [[ST, z, EAo, EAo, EAo, EAo], [z, ST, EAo, EAo, EAo, EAo], [EAo, EAo, ST, z, EAo, EAo],
         [EAo, EAo, z, ST, EAo, EAo], [EAo, EAo, EAo, EAo, ST, z], [EAo, EAo, EAo, EAo, z, ST]],
        [[0, 1, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0], [0, 0, 0, 1, 0, 0],
         [0, 0, 1, 0, 0, 0], [0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 1, 0]])
    T[(1, 0, 0)] = (
        [[z, ST, EAo, EAo, EAo, EAo], [z] * 6, [z, EAo, EAc, z, VAo, VAo],
         [z, EAo, z, EAc, VAo, VAo], [z, EAo, VAo, VAo, EAc, z], [z, EAo, VAo, VAo, z, EAc]],
        [[1, 0, 0, 0, 0, 0], [1] * 6, [1, 0, 0, 1, 0, 0],
         [1, 0, 1, 0, 0, 0], [1, 0, 0, 0, 0, 1], [1, 0, 0, 0, 1, 0]])
    T[(1, 0, 1)] = (
        [[z, EAc, VAo, VAo, z, EAo], [z] * 6, [z, VAo, VAc, z, z, VAo],
         [z, VAo, z, VAc, z, VAo], [z, EAo, VAo, VAo, z, EAc], [z] * 6],
        [[1, 0, 0, 0, 1, 0], [1] * 6, [1, 0, 0, 1, 1, 0],
         [1, 0, 1, 0, 1, 0], [1, 0, 0, 0, 1, 0], [1] * 6])
    T[(0, 0, 1)] = (
        [[EAc, z, VAo, VAo, z, EAo], [z, EAc, VAo, VAo, z, EAo], [VAo, VAo, EAc, z, z, EAo],
         [VAo, VAo, z, EAc, z, EAo], [EAo, EAo, EAo, EAo, z, ST], [z] * 6],
        [[0, 1, 0, 0, 1, 0], [1, 0, 0, 0, 1, 0], [0, 0, 0, 1, 1, 0],
         [0, 0, 1, 0, 1, 0], [0, 0, 0, 0, 1, 0], [1] * 6])
    T[(0, 1, 0)] = (
        [[EAc, z, z, EAo, VAo, VAo], [z, EAc, z, EAo, VAo, VAo], [EAo, EAo, z, ST, EAo, EAo],
         [z] * 6, [VAo, VAo, z, EAo, EAc, z], [VAo, VAo, z, EAo, z, EAc]],
        [[0, 1, 1, 0, 0, 0], [1, 0, 1, 0, 0, 0], [0, 0, 1, 0, 0, 0],
         [1] * 6, [0, 0, 1, 0, 0, 1], [0, 0, 1, 0, 1, 0]])
    T[(1, 1, 0)] = (
        [[z, EAc, z, EAo, VAo, VAo], [z] * 6, [z, EAo, z, EAc, VAo, VAo],
         [z] * 6, [z, VAo, z, VAo, VAc, z], [z, VAo, z, VAo, z, VAc]],
        [[1, 0, 1, 0, 0, 0], [1] * 6, [1, 0, 1, 0, 0, 0],
         [1] * 6, [1, 0, 1, 0, 0, 1], [1, 0, 1, 0, 1, 0]])
    T[(0, 1, 1)] = (
        [[VAc, z, z, VAo, z, VAo], [z, VAc, z, VAo, z, VAo], [VAo, VAo, z, EAc, z, EAo],
         [z] * 6, [VAo, VAo, z, EAo, z, EAc], [z] * 6],
        [[0, 1, 1, 0, 1, 0], [1, 0, 1, 0, 1, 0], [0, 0, 1, 0, 1, 0],
         [1] * 6, [0, 0, 1, 0, 1, 0], [1] * 6])
    T[(1, 1, 1)] = (
        [[z, VAc, z, VAo, z, VAo], [z] * 6, [z, VAo, z, VAc, z, VAo],
         [z] * 6, [z, VAo, z, VAo, z, VAc], [z] * 6],
        [[1, 0, 1, 0, 1, 0], [1] * 6, [1, 0, 1, 0, 1, 0],
         [1] * 6, [1, 0, 1, 0, 1, 0], [1] * 6])
    return {k: (np.array(c, complex), np.array(p, float)) for k, (c, p) in T.items()}


def nearby_correct(m0, I_non_correct, tables):
    """Replace singular entries by the self-term integrals (nearby_correct.m:134-143).

    m0 is the 0-based offset (the reference's m minus one).
    """
    C, P = tables[tuple(m0)]
    return np.where(P == 0.0, C, I_non_correct)


def singular_terms(ko, dx, method='own', ref=None):
    """The five self-term integrals (I_ST, I_EAc, I_EAo, I_VAc, I_VAo).

    'own' uses this package's polar/pyramid quadrature (singular.py);
    'DIRECTFN'/'DEMCEM' call the reference's C++ compiled under oracle/_ref
    (singular_st.m:12-20, singular_ea.m:61-76, singular_va.m:131-147);
    a sequence of five values is used as given (e.g. a fixture's DIRECTFN
    values).
    """
    if isinstance(method, str):
        if method == 'own':
            return singular.all_terms(ko, dx)
        return ref.singular_terms(ko, dx, method)
    vals = tuple(complex(v) for v in method)
    if len(vals) != 5:
        raise ValueError('five self-term values expected')
    return vals


def assembly_nop(L, M, N, ko, dx, method='own', ref=None):
    """Toeplitz generator G_mn (L x M x N x 6) of the N operator (assembly_nop.m:1-142)."""
    if min(L, M, N) < N_MEDIUM:
        raise ValueError("assembly_nop needs every grid side >= 5 (n_medium)")
    R_faces = cube_faces(dx)
    d = np.array([dx, dx, dx])
    mx, my, mz = np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij')
    offs = np.stack([mx.ravel(order='F'), my.ravel(order='F'), mz.ravel(order='F')], axis=1)
    # far: every offset with the 4x4 rule (assembly_nop.m:37-59)
    J = cubature_nop_batch(NP_1D_FAR, offs * d, R_faces, ko, dx)
    J = J.reshape(L, M, N, 6, 6, order='F')
    # medium: offsets [0,5)^3 with 10x10 (assembly_nop.m:67-83)
    a = np.arange(N_MEDIUM)
    m5 = np.array([[i, j, k] for k in a for j in a for i in a])
    Jm = cubature_nop_batch(NP_1D_MEDIUM, m5 * d, R_faces, ko, dx)
    for t, (i, j, k) in enumerate(m5):
        J[i, j, k] = Jm[t]
    # near: [0,2)^3 with 20x20 then singular correction (assembly_nop.m:91-114)
    ST, EAc, EAo, VAc, VAo = singular_terms(ko, dx, method, ref)
    tables = nearby_tables(ST, EAc, EAo, VAc, VAo)
    m2 = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)])
    Jn = cubature_nop_batch(NP_1D_NEAR, m2 * d, R_faces, ko, dx)
    for t, mm in enumerate(m2):
        J[tuple(mm)] = nearby_correct(mm, Jn[t], tables)
    # projection onto the 6 tensor components, 1/(4 pi) (assembly_nop.m:123-136)
    P = vector_projection()
    G = np.einsum('lmnij,ijk->lmnk', J, P)
    return G / (4.0 * np.pi)


def circulant_nop(G, P=None):
    """Embed the Toeplitz octant in a 2L x 2M x 2N circulant (circulant_nop.m:1-43).

    With P = (Px, Py, Pz), Px >= 2L-1, the same embedding in a larger or
    smaller circulant: the block, its signed reflection ending at Px-1, zeros
    between (any such circulant contains the same Toeplitz operator).
    """
    L, M, N, _ = G.shape
    Px, Py, Pz = P or (2 * L, 2 * M, 2 * N)
    C = np.zeros((Px, Py, Pz, 6), complex)
    rL = slice(L - 1, 0, -1)
    rM = slice(M - 1, 0, -1)
    rN = slice(N - 1, 0, -1)
    xL, xM, xN = slice(Px - L + 1, Px), slice(Py - M + 1, Py), slice(Pz - N + 1, Pz)
    C[:L, :M, :N] = G
    C[xL, :M, :N] = G[rL, :, :] * GP_COEFF['L']
    C[:L, xM, :N] = G[:, rM, :] * GP_COEFF['M']
    C[:L, :M, xN] = G[:, :, rN] * GP_COEFF['N']
    C[xL, xM, :N] = G[rL, rM, :] * GP_COEFF['LM']
    C[xL, :M, xN] = G[rL, :, rN] * GP_COEFF['LN']
    C[:L, xM, xN] = G[:, rM, rN] * GP_COEFF['MN']
    C[xL, xM, xN] = G[rL, rM, rN]
    return C


def fft_operator(C, workers=None):
    """Per-component 3D FFT of the circulant (fft_operator.m:6-8)."""
    return sfft.fftn(C, axes=(0, 1, 2), workers=workers)


# ---------------------------------------------------------------- K operator
# Face pairs of the far/medium K cubature and their signs (cubature_kop.cpp:29-36):
# opposite faces of each axis, + for the same side, - for opposite sides.
KOP_OBS = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5])
KOP_SRC = np.array([0, 1, 0, 1, 2, 3, 2, 3, 4, 5, 4, 5])
KOP_SIGN = np.array([1, -1, -1, 1] * 3, float)
# Sign of each K component (x, y, z) in the reflected circulant blocks
# (gperiodic_coeff_kop.m:4-31): component c flips under reflection of axis c.
GP_COEFF_K = {
    'L': np.array([-1, 1, 1], float), 'M': np.array([1, -1, 1], float), 'N': np.array([1, 1, -1], float),
    'LM': np.array([-1, -1, 1], float), 'LN': np.array([-1, 1, -1], float), 'MN': np.array([1, -1, -1], float),
    'LMN': np.array([-1, -1, -1], float),
}
# entries of the 6x6 face-pair matrix that nearby_correct_kop.m:5-77 sets to
# zero (the coincident-panel "ST" entries) for the eight near offsets (0-based)
KOP_ST_ZERO = {(0, 0, 0): [(p, p) for p in range(6)], (1, 0, 0): [(0, 1)], (0, 0, 1): [(4, 5)],
               (0, 1, 0): [(2, 3)], (1, 0, 1): [], (1, 1, 0): [], (0, 1, 1): [], (1, 1, 1): []}


def kop_kernel(ko, R):
    """(1 - exp(-jkR)(1 + jkR)) / R^3 in the form of cubature_kop.cpp:94."""
    return (1.0 - np.exp(-1j * ko * R) * (1.0 + 1j * ko * R)) / R / R / R


def _pair_points(order, i_obs, i_src, R_faces):
    Np, w, u, v = gauss_2d(order)
    p1 = R_faces[:, 0, i_obs]; ap = R_faces[:, 1, i_obs] - p1; bp = R_faces[:, 3, i_obs] - p1
    q1 = R_faces[:, 0, i_src]; aq = R_faces[:, 1, i_src] - q1; bq = R_faces[:, 3, i_src] - q1
    rp = p1[None, :] + np.outer(u, ap) + np.outer(v, bp)
    rq = q1[None, :] + np.outer(u, aq) + np.outer(v, bq)
    return rp, rq, np.outer(w, w)


def _kop_pair_integrals(order, rm, i_obs, i_src, R_faces, ko):
    """sum W G(R) R_c over one face pair for every offset row of rm (K x 3) -> K x 3."""
    rp, rq, W = _pair_points(order, i_obs, i_src, R_faces)
    d = rp[None, :, None, :] + rm[:, None, None, :] - rq[None, None, :, :]     # Rmn = rrp - rrq
    R = np.sqrt(np.sum(d * d, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        G = kop_kernel(ko, R)
    return np.einsum('kpq,kpqc,pq->kc', G, d, W)


def cubature_kop_batch(order, r_m, R_faces, ko, chunk=2048):
    """IK (K x 3) of cubature_kop.cpp:8-125 for many offsets r_m (source at 0):
    sum over the 12 face pairs of sign * sum W G(R) R_c (no dx^4 factor)."""
    r_m = np.atleast_2d(np.asarray(r_m, float))
    out = np.zeros((r_m.shape[0], 3), complex)
    for s0 in range(0, r_m.shape[0], chunk):
        rm = r_m[s0:s0 + chunk]
        for io, iq, sg in zip(KOP_OBS, KOP_SRC, KOP_SIGN):
            out[s0:s0 + chunk] += sg * _kop_pair_integrals(order, rm, io, iq, R_faces, ko)
    return out


def cubature_kop_near(order, m0, R_faces, ko, dx):
    """IK (3,) of cubature_kop_near_sym.m:1-135 for a near offset m0 (0-based, in {0,1}^3):
    the 21 unique face pairs x 3 components, completion, the coincident-panel
    zeroing of nearby_correct_kop.m, then the row combination of :120-131."""
    rm = (np.asarray(m0, float) * dx)[None, :]
    I = np.zeros((6, 6, 3), complex)
    for io, iq in zip(INDEX_OBS, INDEX_SRC):
        if (io, iq) in KOP_ST_ZERO[tuple(m0)]:
            continue                     # zeroed by the nearby correction (and singular)
        I[io, iq] = _kop_pair_integrals(order, rm, io, iq, R_faces, ko)[0]
    for (dr, dc), (sr, sc) in SYM_FILL:
        I[dr, dc] = I[sr, sc]
    for (p, q) in KOP_ST_ZERO[tuple(m0)]:
        I[p, q] = 0.0
    sgn = np.array([-1, 1, -1, 1, -1, 1], float)
    comp = [0, 0, 1, 1, 2, 2]
    J = np.stack([sgn[r] * I[r, :, comp[r]] for r in range(6)])     # 6 x 6
    return np.array([np.sum(J[:, 1] - J[:, 0]), np.sum(J[:, 3] - J[:, 2]), np.sum(J[:, 5] - J[:, 4])])


def assembly_kop(L, M, N, ko, dx):
    """Toeplitz generator IK_mn (L x M x N x 3) of the K operator (assembly_kop.m:1-121)."""
    if min(L, M, N) < N_MEDIUM:
        raise ValueError("assembly_kop needs every grid side >= 5 (n_medium)")
    R_faces = cube_faces(dx)
    d = np.array([dx, dx, dx])
    mx, my, mz = np.meshgrid(np.arange(L), np.arange(M), np.arange(N), indexing='ij')
    offs = np.stack([mx.ravel(order='F'), my.ravel(order='F'), mz.ravel(order='F')], axis=1)
    IK = cubature_kop_batch(NP_1D_FAR, offs * d, R_faces, ko).reshape(L, M, N, 3, order='F')
    mm = np.arange(N_MEDIUM)
    mx, my, mz = np.meshgrid(mm, mm, mm, indexing='ij')
    offs = np.stack([mx.ravel(order='F'), my.ravel(order='F'), mz.ravel(order='F')], axis=1)
    IK[:N_MEDIUM, :N_MEDIUM, :N_MEDIUM] = cubature_kop_batch(NP_1D_MEDIUM, offs * d, R_faces, ko).reshape(
        N_MEDIUM, N_MEDIUM, N_MEDIUM, 3, order='F')
    for m0 in [(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]:
        IK[m0] = cubature_kop_near(NP_1D_NEAR, m0, R_faces, ko, dx)
    return IK * (dx ** 4 / (1j * ko) ** 2 / (4.0 * np.pi))


def circulant_kop(G, P=None):
    """Embed the K Toeplitz octant in the circulant (circulant_kop.m:22-39);
    P = (Px, Py, Pz) with Px >= 2L-1 embeds it in another circulant size."""
    L, M, N, _ = G.shape
    Px, Py, Pz = P or (2 * L, 2 * M, 2 * N)
    C = np.zeros((Px, Py, Pz, 3), complex)
    rL, rM, rN = slice(L - 1, 0, -1), slice(M - 1, 0, -1), slice(N - 1, 0, -1)
    xL, xM, xN = slice(Px - L + 1, Px), slice(Py - M + 1, Py), slice(Pz - N + 1, Pz)
    C[:L, :M, :N] = G
    C[xL, :M, :N] = G[rL, :, :] * GP_COEFF_K['L']
    C[:L, xM, :N] = G[:, rM, :] * GP_COEFF_K['M']
    C[:L, :M, xN] = G[:, :, rN] * GP_COEFF_K['N']
    C[xL, xM, :N] = G[rL, rM, :] * GP_COEFF_K['LM']
    C[xL, :M, xN] = G[rL, :, rN] * GP_COEFF_K['LN']
    C[:L, xM, xN] = G[:, rM, rN] * GP_COEFF_K['MN']
    C[xL, xM, xN] = G[rL, rM, rN] * GP_COEFF_K['LMN']
    return C


def get_operators(L, M, N, dx, f, op='N', type_='FFTofC', method='own', ref=None):
    """getOPERATORS.m:1-127 (ko from :48, dispatch 'N' :74-96, 'K' :101-115)."""
    ko = 2.0 * np.pi / C0 * f
    if op == 'N':
        G = assembly_nop(L, M, N, ko, dx, method, ref)
        circ = circulant_nop
    elif op == 'K':
        G = assembly_kop(L, M, N, ko, dx)
        circ = circulant_kop
    else:
        raise NotImplementedError("oracle covers the N and K operators")
    if type_ == 'T':
        return G
    C = circ(G)
    if type_ == 'C':
        return C
    return fft_operator(C)
