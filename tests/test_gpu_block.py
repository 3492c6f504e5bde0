"""Block GMRES (several excitations of one body) on the device.

The reference solves excitation ports one pgmres at a time
(VSIE_Solver.m:387-407); the block solver must reach, for every column, the
solution of that column's own solve: checked against oracle.krylov.pgmres
run per column to 1e-11 (the two solutions then agree to the conditioning
times the tolerance), with each column's true residual <= tol.
"""
import numpy as np
import pytest

from oracle import fields, krylov, mvp
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _problem(n=12, dx=6.25e-3, f=298.2e6, eps=50.0, sig=0.5):
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    inside = X ** 2 + 0.8 * Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2
    em = fields.em_constants(f)
    e_r = np.where(inside, eps - 1j * sig / (em['eo'] * em['omega']), 1.0 + 0j)
    ko = em['ko']
    waves = [([0, 0, ko], [1, 0, 0]), ([0, 0, ko], [0, 1, 0]), ([ko, 0, 0], [0, 0, 1]), ([0, -ko, 0], [1, 0, 0])]
    rhs = []
    for k, pol in waves:
        Einc, _ = fields.plane_wave_excitation(r, k, em['omega_mu'], pol)
        Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
        rhs.append(b)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    return fN, Mr, Mc, Gram, idx, np.stack(rhs, axis=1)


@pytest.fixture(scope='module')
def problem():
    return _problem()


def _op(problem, prec=None):
    from marie_amd import lib
    fN, Mr, Mc, Gram, idx, B = problem
    n = fN.shape[0] // 2
    op = lib.Operator(n, n, n, precisions=prec or lib.PREC_F64)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    return op


def _reference(problem):
    fN, Mr, Mc, Gram, idx, B = problem
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    return np.stack([krylov.pgmres(A, B[:, q], 50, 1e-11, 200)[0] for q in range(B.shape[1])], axis=1)


@pytest.fixture(scope='module')
def reference(problem):
    return _reference(problem)


@pytest.mark.parametrize("restart", [10, 3])
def test_block_matches_per_column_solves(problem, reference, restart):
    print("case block fp64 restart", restart, flush=True)
    op = _op(problem)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B, restart=restart, tol=1e-10, maxit=100)
    assert np.all(flags == 0) and np.all(relres <= 1e-10), relres
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-8, q
    # true residual history ends with the final worst column
    assert abs(rv[-1] - relres.max()) <= 1e-12


def test_block_two_pass_gram_schmidt(problem, reference):
    print("case block cgs2", flush=True)
    op = _op(problem)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-10, maxit=100, cgs_passes=2)
    assert np.all(flags == 0) and np.all(relres <= 1e-10), relres
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-8, q


def test_block_single_column(problem, reference):
    print("case block single column", flush=True)
    op = _op(problem)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B[:, 2], restart=20, tol=1e-10, maxit=100)
    assert flags[0] == 0 and relres[0] <= 1e-10
    assert _rel(X[:, 0], reference[:, 2]) <= 1e-8


def test_block_complex64_products(problem, reference):
    """mode 1: complex64 products, FP64 basis and residual replacement."""
    from marie_amd import lib
    print("case block mode 1", flush=True)
    op = _op(problem, lib.PREC_F64 | lib.PREC_F32)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-8, maxit=100, mode=1, inner_tol=1e-5)
    assert np.all(flags == 0) and np.all(relres <= 1e-8)
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-6


def test_block_graph_replay_equals_eager(problem):
    """Block steps replayed from captured HIP graphs give the eager results
    exactly (restart 3 forces several cycles, so graphs are replayed)."""
    print("case block graphs", flush=True)
    op = _op(problem)
    B = problem[-1]
    Xe, fe, re, ite, rve = op.block_gmres(B, restart=3, tol=1e-9, maxit=100, use_graph=False)
    Xg, fg, rg, itg, rvg = op.block_gmres(B, restart=3, tol=1e-9, maxit=100, use_graph=True)
    assert ite[1] >= 3 and itg == ite
    assert np.array_equal(Xg, Xe)
    assert np.array_equal(rvg, rve)


def test_block_dependent_columns(problem, reference):
    """Repeated and proportional right-hand sides (a rank-deficient block)."""
    print("case block dependent columns", flush=True)
    op = _op(problem)
    b = problem[-1][:, 0]
    B = np.stack([b, 2j * b, b], axis=1)
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-10, maxit=100)
    assert np.all(flags == 0)
    assert _rel(X[:, 0], reference[:, 0]) <= 1e-8
    assert _rel(X[:, 1], 2j * reference[:, 0]) <= 1e-8
    assert _rel(X[:, 2], reference[:, 0]) <= 1e-8


def test_block_zero_column_rejected(problem):
    from marie_amd import lib
    op = _op(problem)
    B = problem[-1][:, :2].copy()
    B[:, 1] = 0
    with pytest.raises(lib.MarieError):
        op.block_gmres(B)


def test_block_batch_of_uneven_problems():
    """Several bodies / frequencies (separate handles, uneven sizes) solved
    concurrently give each problem's own block solution; a handle used twice
    is rejected."""
    import torch
    from marie_amd import lib
    print("case block batch", flush=True)
    probs = [_problem(10, 6.25e-3, 298.2e6), _problem(14, 5e-3, 128e6, eps=40.0, sig=0.3),
             _problem(12, 6.25e-3, 200e6, eps=60.0, sig=0.8)]
    ops, Bs, Xs, nrhs, refs = [], [], [], [], []
    for fN, Mr, Mc, Gram, idx, B in probs:
        n = Mr.shape[0]
        op = lib.Operator(n, n, n)
        op.set_fN(fN)
        op.set_material(Mr, Mc, Gram)
        op.set_index(idx)
        ops.append(op)
        Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to('cuda:0')
        Bs.append(Bd)
        Xs.append(torch.zeros_like(Bd))
        nrhs.append(B.shape[1])
        Xr, fl, rr, it, rv = op.block_gmres(B, restart=8, tol=1e-9, maxit=100)
        assert np.all(fl == 0)
        refs.append(Xr)
    torch.cuda.synchronize()
    flags, relres, iters = lib.block_gmres_batch(ops, [b.data_ptr() for b in Bs], [x.data_ptr() for x in Xs], nrhs,
                                                 restart=8, tol=1e-9, maxit=100)
    for i in range(len(probs)):
        assert np.all(flags[i] == 0) and np.all(relres[i] <= 1e-9)
        X = Xs[i].cpu().numpy().T
        assert _rel(X, refs[i]) <= 1e-12          # the same deterministic solve
        fN, Mr, Mc, Gram, idx, B = probs[i]
        for q in range(B.shape[1]):
            r = B[:, q] - mvp.mv_AN(X[:, q], fN, Mr, Mc, Gram, 'notransp', idx)
            assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, q])
    with pytest.raises(lib.MarieError):
        lib.block_gmres_batch([ops[0], ops[0]], [Bs[0].data_ptr()] * 2, [Xs[0].data_ptr()] * 2, [nrhs[0]] * 2)


@pytest.mark.parametrize("mode,restart", [(0, 50), (3, 40)])
def test_block_columns_option(problem, reference, mode, restart):
    """MARIE_BLOCK_COLUMNS: each column one GMRES solve in `mode` (0: pgmres,
    3: block-orthogonalised complex64 cycles with FP64 refinement), on one
    handle and through the batch entry over two handles; every column's true
    residual by the oracle's product <= tol and its solution within the tolerance
    of the oracle's per-column pgmres."""
    import torch
    from marie_amd import lib
    print("case block columns mode", mode, flush=True)
    fN, Mr, Mc, Gram, idx, B = problem
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    dev = torch.device('cuda:0')
    ops = [_op(problem, lib.PREC_F64 | lib.PREC_F32) for _ in range(2)]
    Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    Xd = torch.zeros_like(Bd)
    fl, rr, it, rv = ops[0].block_gmres_device(Bd.data_ptr(), Xd.data_ptr(), B.shape[1], restart=restart, tol=1e-6,
                                               maxit=200, mode=mode, inner_tol=1e-7, columns=True)
    X = Xd.cpu().numpy().T
    assert np.all(fl == 0) and np.all(rr <= 1e-6) and it[1] >= B.shape[1]
    for q in range(B.shape[1]):
        assert _rel(A(X[:, q]), B[:, q]) <= 1e-6
        assert _rel(X[:, q], reference[:, q]) <= 1e-4
    Xs = [torch.zeros_like(Bd), torch.zeros_like(Bd)]
    fls, rrs, its = lib.block_gmres_batch(ops, [Bd.data_ptr(), Bd.data_ptr()], [x.data_ptr() for x in Xs],
                                          [B.shape[1]] * 2, restart=restart, tol=1e-6, maxit=200, mode=mode,
                                          inner_tol=1e-7, columns=True)
    for i in range(2):
        assert np.all(fls[i] == 0) and np.all(rrs[i] <= 1e-6)
        assert torch.equal(Xs[i], Xd)      # same kernels, same order, same stream semantics: identical
    for o in ops:
        o.close()
