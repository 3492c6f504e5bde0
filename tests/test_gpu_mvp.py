"""GPU parity of the FFT-circulant matvecs against the oracle (mv_AN.m / mv_N.m / mv_L.m).

Every case calls the HIP library through its C-ABI (marie_amd.lib) and
compares with oracle.mvp on the same seeded inputs. Tolerances: complex128
<= 1e-12 relative (2-norm over the output vector), complex64 <= 2e-6.
"""
import numpy as np
import pytest

from oracle import mvp
from oracle import operators as oop

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def _case(L, M, N, P=None, frac=0.6, seed=0):
    rng = np.random.default_rng(seed)
    P = P or (2 * L, 2 * M, 2 * N)
    fN = rng.standard_normal(P + (6,)) + 1j * rng.standard_normal(P + (6,))
    Mr = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    Mc = rng.standard_normal((L, M, N)) + 1j * rng.standard_normal((L, M, N))
    n3 = 3 * L * M * N
    idx = np.sort(rng.choice(n3, max(1, int(frac * n3)), replace=False))
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    return fN, Mr, Mc, idx, x


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _op(L, M, N, P, fN, Mr, Mc, Gram, idx, prec=None):
    from marie_amd import lib as L_
    op = L_.Operator(L, M, N, P, precisions=prec or L_.PREC_F64)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    return op


SHAPES = [
    ((8, 8, 8), None),            # powers of two
    ((5, 6, 7), None),            # 10, 12, 14: radices 2, 3, 4, 5, 7
    ((12, 10, 9), None),          # 24, 20, 18
    ((33, 33, 33), None),         # 66 = 2*3*11 (config-1 grid of generatedomain)
    ((16, 12, 20), (40, 24, 48)), # circulant larger than 2L (fftn with explicit size)
    ((32, 32, 32), None),
]


@pytest.mark.parametrize("shape,P", SHAPES)
@pytest.mark.parametrize("transp", [False, True])
def test_mv_AN_parity_f64(shape, P, transp):
    L, M, N = shape
    P = P or (2 * L, 2 * M, 2 * N)
    print("case mv_AN f64", shape, P, "transp", transp, flush=True)
    fN, Mr, Mc, idx, x = _case(L, M, N, P)
    Gram = 0.37
    op = _op(L, M, N, P, fN, Mr, Mc, Gram, idx)
    from marie_amd import lib as L_
    y = op.apply(x, L_.OP_AN, transp)
    yr = mvp.mv_AN(x, fN, Mr, Mc, Gram, 'transp' if transp else 'notransp', idx)
    assert _rel(y, yr) <= TOL64


@pytest.mark.parametrize("shape,P", SHAPES[:4])
def test_mv_AN_parity_f32(shape, P):
    L, M, N = shape
    P = P or (2 * L, 2 * M, 2 * N)
    print("case mv_AN f32", shape, P, flush=True)
    fN, Mr, Mc, idx, x = _case(L, M, N, P, seed=3)
    from marie_amd import lib as L_
    op = _op(L, M, N, P, fN, Mr, Mc, 0.5, idx, prec=L_.PREC_F32)
    y = op.apply(x, L_.OP_AN, False)
    yr = mvp.mv_AN(x, fN, Mr, Mc, 0.5, 'notransp', idx)
    assert _rel(y, yr) <= TOL32


@pytest.mark.parametrize("opname", ["N", "L"])
def test_mv_N_L_parity(opname):
    L, M, N = 9, 8, 10
    print("case mv_" + opname, flush=True)
    fN, Mr, Mc, idx, x = _case(L, M, N, frac=1.0, seed=5)
    idx = np.arange(3 * L * M * N)      # E_field_Nop.m:69 uses every DOF
    x = np.random.default_rng(6).standard_normal(idx.size) + 0j
    from marie_amd import lib as L_
    op = _op(L, M, N, None, fN, Mr, Mc, 0.25, idx)
    if opname == "N":
        y = op.apply(x, L_.OP_N)
        yr = mvp.mv_N(x, fN, L, M, N, idx)
    else:
        y = op.apply(x, L_.OP_L)
        yr = mvp.mv_L(x, fN, L, M, N, 0.25, idx)
    assert _rel(y, yr) <= TOL64


def test_multi_rhs_matches_single():
    L, M, N = 10, 12, 8
    print("case nrhs=3", flush=True)
    fN, Mr, Mc, idx, _ = _case(L, M, N, seed=7)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((idx.size, 3)) + 1j * rng.standard_normal((idx.size, 3))
    op = _op(L, M, N, None, fN, Mr, Mc, 1.0, idx)
    Y = op.apply(X)
    for k in range(3):
        yr = mvp.mv_AN(X[:, k], fN, Mr, Mc, 1.0, 'notransp', idx)
        assert _rel(Y[:, k], yr) <= TOL64


@pytest.mark.parametrize("transp", [False, True])
def test_multi_rhs_column_groups(transp, monkeypatch):
    """A multi-column product whose x and y passes run in groups of columns
    (MARIE_XY_GROUP=2 over 5 columns: groups of 2, 2, 1, the path large
    grids take to keep each group's buffers in the last-level cache) equals
    the oracle per column and, bit for bit, the one-group product, in FP64,
    complex64 and through the allocation-free entry point."""
    import torch
    from marie_amd import lib as L_
    L, M, N, nr = 10, 12, 8, 5
    print("case nrhs=5 in column groups of 2, transp", transp, flush=True)
    fN, Mr, Mc, idx, _ = _case(L, M, N, seed=17)
    rng = np.random.default_rng(18)
    X = rng.standard_normal((nr, idx.size)) + 1j * rng.standard_normal((nr, idx.size))
    prec = L_.PREC_F64 | L_.PREC_F32
    monkeypatch.setenv('MARIE_XY_GROUP', '2')
    grouped = _op(L, M, N, None, fN, Mr, Mc, 1.5, idx, prec)
    monkeypatch.setenv('MARIE_XY_GROUP', '0')
    whole = _op(L, M, N, None, fN, Mr, Mc, 1.5, idx, prec)
    dev = torch.device('cuda:0')
    for p, dt in ((L_.PREC_F64, torch.complex128), (L_.PREC_F32, torch.complex64)):
        Xd = torch.tensor(X, dtype=dt, device=dev)
        out = {}
        for name, op in (('grouped', grouped), ('whole', whole)):
            Y = torch.empty_like(Xd)
            op.apply_device(Xd.data_ptr(), Y.data_ptr(), transp=transp, prec=p, nrhs=nr)
            torch.cuda.synchronize()
            out[name] = Y.cpu().numpy()
        assert np.array_equal(out['grouped'], out['whole'])
        Yn = torch.empty_like(Xd)
        grouped.apply_device_noalloc(Xd.data_ptr(), Yn.data_ptr(), transp=transp, prec=p, nrhs=nr)
        torch.cuda.synchronize()
        assert np.array_equal(Yn.cpu().numpy(), out['whole'])
        tol = TOL64 if p == L_.PREC_F64 else TOL32
        for k in range(nr):
            yr = mvp.mv_AN(X[k], fN, Mr, Mc, 1.5, 'transp' if transp else 'notransp', idx)
            assert _rel(out['grouped'][k], yr) <= tol, (p, k)
    grouped.close()
    whole.close()


def test_unsorted_one_based_index():
    L, M, N = 6, 7, 5
    print("case unsorted 1-based idx", flush=True)
    fN, Mr, Mc, idx, x = _case(L, M, N, seed=9)
    perm = np.random.default_rng(10).permutation(idx.size)
    idx_u = idx[perm]
    from marie_amd import lib as L_
    op = L_.Operator(L, M, N)
    op.set_fN(fN)
    op.set_material(Mr, Mc, 2.0)
    op.set_index(idx_u + 1, one_based=True)
    y = op.apply(x[perm])
    yr = mvp.mv_AN(x[perm], fN, Mr, Mc, 2.0, 'notransp', idx_u)
    assert _rel(y, yr) <= TOL64


def test_physical_operator_symmetry():
    """u^T N v == v^T N u for an assembled-like (parity-consistent) operator."""
    L, M, N = 7, 6, 5
    print("case complex symmetry", flush=True)
    rng = np.random.default_rng(11)
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    fN = oop.fft_operator(oop.circulant_nop(G))
    idx = np.arange(3 * L * M * N)
    from marie_amd import lib as L_
    op = _op(L, M, N, None, fN, np.ones((L, M, N)), np.zeros((L, M, N)), 1.0, idx)
    u = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    v = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    a = u @ op.apply(v, L_.OP_N)
    b = v @ op.apply(u, L_.OP_N)
    assert abs(a - b) <= 1e-12 * abs(a)
    # against the brute-force Toeplitz sum
    y = op.apply(v, L_.OP_N)
    yr = -mvp.toeplitz_apply_dense(v, G, np.zeros((L, M, N)), np.ones((L, M, N)), 0.0, idx)
    assert _rel(y, yr) <= TOL64


def test_error_paths():
    from marie_amd import lib as L_
    print("case error paths", flush=True)
    op = L_.Operator(4, 4, 4)
    with pytest.raises(L_.MarieError):
        op.set_index(np.array([0, 0]))           # duplicate
    with pytest.raises(L_.MarieError):
        op.set_index(np.array([3 * 64]))         # out of range
    with pytest.raises(L_.MarieError):
        op.apply(np.ones(0))                     # operator and index not set
    op.set_index(np.array([0, 5, 9], np.int64))
    with pytest.raises(ValueError):
        op.apply(np.ones(2))                     # x shorter than the index: refused before the C call
    with pytest.raises(ValueError):
        op.gmres(np.ones(4))
    op.set_index(np.array([], np.int64))
    fN = np.ones((8, 8, 8, 6), complex)
    op.set_fN(fN)
    op.set_material(np.ones((4, 4, 4)), np.zeros((4, 4, 4)), 1.0)
    assert op.apply(np.zeros(0, complex)).size == 0  # empty index: no-op
    bad = fN.copy()
    bad[0, 0, 0, 0] = np.nan
    with pytest.raises(L_.MarieError):
        op.set_fN(bad)
    with pytest.raises(L_.MarieError):
        L_.Operator(4, 4, 4, (8, 8, 34))         # 34 = 2 * 17: prime factor above 13


@pytest.mark.parametrize("prec,variants,shape", [
    ("f64", [4, 8, 16, -4, -8, -16], (12, 10, 16)),   # LDS engine (+ register engine on the 2^k side)
    ("f32", [8, 16, 32, -4, -8, -16], (12, 10, 16)),
    ("f64", [-4, -8, -16, 8], (16, 8, 32)),          # register engine on every pass
    ("f32", [-4, -8, -16, 16], (16, 8, 32)),
])
def test_every_tile_variant(prec, variants, shape):
    """Each compiled tile shape gives the oracle's product (both fN storage forms).

    Positive values select the LDS-stage engine (elements per thread), negative
    values the register engine (points per thread) where the side is 2^k."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case tile variants", prec, flush=True)
    rng = np.random.default_rng(21)
    # parity-consistent operator -> octant storage; random operator -> full storage
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    fN_sym = oop.fft_operator(oop.circulant_nop(G))
    _, Mr, Mc, idx, x = _case(L, M, N, seed=22)
    fN_rand = _case(L, M, N, seed=23)[0]
    p = L_.PREC_F64 if prec == "f64" else L_.PREC_F32
    tol = TOL64 if prec == "f64" else TOL32
    for fN in (fN_sym, fN_rand):
        op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=p)
        yr = mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'notransp', idx)
        for exy in variants:
            for ez in variants:
              for zg in (0, 1):
                op.set_tiles([exy, exy, ez, exy, exy, zg], p)
                assert _rel(op.apply(x), yr) <= tol, (exy, ez, zg)
                assert _rel(op.apply(x, L_.OP_AN, True),
                            mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'transp', idx)) <= tol, (exy, ez, zg)
        # round trip of the stored operator (octant expand or full copy)
        if prec == "f64":
            assert _rel(op.get_fN(), fN) <= 1e-15


@pytest.mark.parametrize("prec,shape", [("f64", (16, 8, 32)), ("f32", (16, 8, 32)), ("f64", (12, 10, 16)),
                                        ("f64", (5, 6, 64))])
def test_z_split_variants(prec, shape):
    """The z kernel with one thread group per component (codes 100*TX + E)
    gives the oracle's products: N (both storage forms, both directions, two
    right-hand sides) and K."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case z split", prec, shape, flush=True)
    p = L_.PREC_F64 if prec == "f64" else L_.PREC_F32
    tol = TOL64 if prec == "f64" else TOL32
    rng = np.random.default_rng(31)
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    fN_sym = oop.fft_operator(oop.circulant_nop(G))
    _, Mr, Mc, idx, x = _case(L, M, N, seed=32)
    x2 = np.stack([x, rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)], axis=1)
    fN_rand = _case(L, M, N, seed=33)[0]
    codes = [408, 404, 416, 808]
    for fN in (fN_sym, fN_rand):
        op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=p)
        yr = [mvp.mv_AN(x2[:, q], fN, Mr, Mc, 0.8, 'notransp', idx) for q in range(2)]
        yt = mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'transp', idx)
        for code in codes:
            op.set_tiles([-8, -8, -code, -8, -8, 0], p)
            y = op.apply(x2)
            for q in range(2):
                assert _rel(y[:, q], yr[q]) <= tol, (code, q)
            assert _rel(op.apply(x, L_.OP_AN, True), yt) <= tol, code
    # K: cross product, octant storage of a physical fK
    ko, dx = 2 * np.pi / 299792458.0 * 298e6, 6.25e-3
    fK = oop.fft_operator(oop.circulant_kop(oop.assembly_kop(L, M, N, ko, dx)))
    op = L_.Operator(L, M, N, precisions=p)
    op.set_fK(fK)
    op.set_index(idx)
    yk = mvp.mv_K(x, fK, L, M, N, idx)
    for code in codes:
        op.set_tiles([-8, -8, -code, -8, -8, 0], p)
        assert _rel(op.apply(x, L_.OP_K), yk) <= tol, ('K', code)


@pytest.mark.parametrize("shape", [(16, 16, 32), (20, 64, 12), (24, 128, 8)])
def test_column_pair_y_pass(shape):
    """complex64 y passes on column pairs (codes 200 + E: 16-byte accesses,
    two lines per thread) give the oracle's products: both directions, two
    right-hand sides, both fN storage forms; power-of-two y sides with an even
    x side that is not always a power of two (40, 48)."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case column pairs", shape, flush=True)
    rng = np.random.default_rng(41)
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    fN_sym = oop.fft_operator(oop.circulant_nop(G))
    _, Mr, Mc, idx, x = _case(L, M, N, seed=42)
    x2 = np.stack([x, rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)], axis=1)
    fN_rand = _case(L, M, N, seed=43)[0]
    for fN in (fN_sym, fN_rand):
        op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=L_.PREC_F32)
        yr = [mvp.mv_AN(x2[:, q], fN, Mr, Mc, 0.8, 'notransp', idx) for q in range(2)]
        yt = mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'transp', idx)
        for code in (204, 208, 216):
            for ex in (-8, 16):
                op.set_tiles([ex, -code, 8, -code, ex, 0], L_.PREC_F32)
                y = op.apply(x2)
                for q in range(2):
                    assert _rel(y[:, q], yr[q]) <= TOL32, (code, ex, q)
                assert _rel(op.apply(x, L_.OP_AN, True), yt) <= TOL32, (code, ex)


@pytest.mark.parametrize("shape", [(16, 16, 32), (20, 64, 12), (24, 128, 8)])
def test_two_line_y_pass_fp64(shape):
    """FP64 y passes on two lines per thread (codes 200 + E: columns TX apart,
    each load a contiguous run) give the oracle's products to 1e-12: both
    directions, two right-hand sides; x sides 32, 40, 48 leave the last
    2 TX-column tile partly empty."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case fp64 two-line y", shape, flush=True)
    rng = np.random.default_rng(45)
    fN, Mr, Mc, idx, x = _case(L, M, N, seed=46)
    x2 = np.stack([x, rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)], axis=1)
    op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=L_.PREC_F64)
    yr = [mvp.mv_AN(x2[:, q], fN, Mr, Mc, 0.8, 'notransp', idx) for q in range(2)]
    yt = mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'transp', idx)
    for code in (204, 208, 216):
        for fwd, inv in ((-code, -8), (-8, -code), (-code, -code)):
            op.set_tiles([-8, fwd, -8, inv, -4, 0], L_.PREC_F64)
            y = op.apply(x2)
            for q in range(2):
                assert _rel(y[:, q], yr[q]) <= TOL64, (code, fwd, inv, q)
            assert _rel(op.apply(x, L_.OP_AN, True), yt) <= TOL64, (code, fwd, inv)


MIXED = [
    ((112, 96, 10), ([-14, -12, 8, -12, -14], [-28, -24, 8, -24, -28], [-28, -12, 8, -24, -14])),  # 224, 192
    ((48, 56, 10), ([-12, -14, 8, -14, -12], [-24, -28, 8, -28, -24], [-24, -14, 8, -28, -12])),    # 96, 112
    ((80, 48, 10), ([-10, -12, 8, -12, -10], [-20, -24, 8, -24, -20])),                             # 160, 96
]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape,codeset", MIXED)
def test_mixed_radix_register_engine(prec, shape, codeset):
    """Register-engine x and y kernels on mixed-radix sides (96, 192 = 2^k * 3
    with 12 or 24 points per thread; 112, 224 = 2^k * 7 with 14 or 28; 160 =
    2^5 * 5 with 10 or 20) give the oracle's products, both directions; the z
    side (20) runs on the LDS-stage engine."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case mixed radix register engine", prec, shape, flush=True)
    p = L_.PREC_F64 if prec == "f64" else L_.PREC_F32
    tol = TOL64 if prec == "f64" else TOL32
    fN, Mr, Mc, idx, x = _case(L, M, N, frac=0.3, seed=51)
    op = _op(L, M, N, None, fN, Mr, Mc, 0.7, idx, prec=p)
    assert op.P == (2 * L, 2 * M, 2 * N)
    yr = mvp.mv_AN(x, fN, Mr, Mc, 0.7, 'notransp', idx)
    yt = mvp.mv_AN(x, fN, Mr, Mc, 0.7, 'transp', idx)
    for codes in codeset:
        op.set_tiles(codes + [0], p)
        assert _rel(op.apply(x), yr) <= tol, codes
        assert _rel(op.apply(x, L_.OP_AN, True), yt) <= tol, codes


@pytest.mark.parametrize("prec,shape,zcode", [
    ("f64", (10, 12, 80), 10),      # z side 160 = 2^5 * 5
    ("f64", (10, 12, 96), 12),      # z side 192 = 2^6 * 3
    ("f32", (10, 12, 112), 14),     # z side 224 = 2^5 * 7 (complex64 only)
    ("f32", (10, 12, 48), 12),      # z side 96
    ("f64", (10, 12, 112), 414),    # z side 224: one thread group per component, 14 points, 4 columns
    ("f32", (10, 12, 112), 414),
])
def test_mixed_radix_z(prec, shape, zcode):
    """The register-engine z kernel on mixed-radix sides gives the oracle's N
    products (two right-hand sides, both directions) and K product; an untuned
    handle takes the same plan by default."""
    from marie_amd import lib as L_
    L, M, N = shape
    print("case mixed radix z", prec, shape, zcode, flush=True)
    p = L_.PREC_F64 if prec == "f64" else L_.PREC_F32
    tol = TOL64 if prec == "f64" else TOL32
    rng = np.random.default_rng(61)
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    for c in range(6):
        if oop.GP_COEFF['L'][c] < 0:
            G[0, :, :, c] = 0
        if oop.GP_COEFF['M'][c] < 0:
            G[:, 0, :, c] = 0
        if oop.GP_COEFF['N'][c] < 0:
            G[:, :, 0, c] = 0
    fN = oop.fft_operator(oop.circulant_nop(G))
    _, Mr, Mc, idx, x = _case(L, M, N, frac=0.5, seed=62)
    x2 = np.stack([x, rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)], axis=1)
    op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=p)
    yr = [mvp.mv_AN(x2[:, q], fN, Mr, Mc, 0.8, 'notransp', idx) for q in range(2)]
    yt = mvp.mv_AN(x, fN, Mr, Mc, 0.8, 'transp', idx)
    for cfg in ([8, 8, -zcode, 8, 8, 0], None):
        if cfg is not None:
            op.set_tiles(cfg, p)
        y = op.apply(x2)
        for q in range(2):
            assert _rel(y[:, q], yr[q]) <= tol, (cfg, q)
        assert _rel(op.apply(x, L_.OP_AN, True), yt) <= tol, cfg
        op = _op(L, M, N, None, fN, Mr, Mc, 0.8, idx, prec=p)     # a fresh, untuned handle next
    ko, dx = 2 * np.pi / 299792458.0 * 298e6, 6.25e-3
    fK = oop.fft_operator(oop.circulant_kop(oop.assembly_kop(L, M, N, ko, dx)))
    opk = L_.Operator(L, M, N, precisions=p)
    opk.set_fK(fK)
    opk.set_index(idx)
    opk.set_tiles([8, 8, -zcode, 8, 8, 0], p)
    assert _rel(opk.apply(x, L_.OP_K), mvp.mv_K(x, fK, L, M, N, idx)) <= tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tile_request_between_variants_and_invalid_codes(prec):
    """A positive tile request above the compiled variants (32 elements per
    thread in FP64 on a 512-point side) runs the largest compiled variant with
    a tile that fits it, and gives the oracle's product; codes that name no
    compiled variant are refused."""
    from marie_amd import lib as L_
    L, M, N = 256, 6, 5
    print("case tile request between variants", prec, flush=True)
    p = L_.PREC_F64 if prec == "f64" else L_.PREC_F32
    tol = TOL64 if prec == "f64" else TOL32
    fN, Mr, Mc, idx, x = _case(L, M, N, frac=0.4, seed=71)
    op = _op(L, M, N, None, fN, Mr, Mc, 0.9, idx, prec=p)
    assert op.P == (512, 12, 10)
    yr = mvp.mv_AN(x, fN, Mr, Mc, 0.9, 'notransp', idx)
    for cfg in ([32, 32, 16, 32, 32], [16, 16, 8, 16, 16], [4, 8, 4, 8, 4]):
        op.set_tiles(cfg, p)
        assert _rel(op.apply(x), yr) <= tol, cfg
    for bad in ([3, 8, 8, 8, 8], [8, 8, -414 - 1, 8, 8], [8, -404, 8, 8, 8], [0, 8, 8, 8, 8]):
        with pytest.raises(L_.MarieError):
            op.set_tiles(bad, p)
