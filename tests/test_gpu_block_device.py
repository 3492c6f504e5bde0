"""Block GMRES with device-driven block steps (MARIE_BLOCK_DEVICE): the
product, block Gram-Schmidt, CholeskyQR and the Householder least-squares
update all on the device, each step optionally one HIP graph replayed every
cycle, and a batch of uneven problems driven from one host thread.

The arithmetic is that of the host-driven block solver, so every column must
reach the solution of its own pgmres solve (oracle.krylov.pgmres per column
to 1e-11, VSIE_Solver.m:387-407 solves ports one at a time), the true
residual by the oracle's product must meet tol (pgmres.m:175), and the
host-driven solver's solution must agree to rounding.
"""
import numpy as np
import pytest

from oracle import mvp

from test_gpu_block import _op, _problem, _rel, _reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def problem():
    return _problem()


@pytest.fixture(scope='module')
def reference(problem):
    return _reference(problem)


def _stream_op(problem, prec=None):
    import torch
    op = _op(problem, prec)
    st = torch.cuda.Stream(device='cuda:0')
    op.set_stream(st.cuda_stream)
    return op, st


def _true_res(problem, X):
    fN, Mr, Mc, Gram, idx, B = problem
    return [np.linalg.norm(B[:, q] - mvp.mv_AN(X[:, q], fN, Mr, Mc, Gram, 'notransp', idx)) /
            np.linalg.norm(B[:, q]) for q in range(B.shape[1])]


@pytest.mark.parametrize("restart,use_graph", [(10, False), (10, True), (3, True)])
def test_device_steps_match_per_column_solves(problem, reference, restart, use_graph):
    print("case device-driven block restart", restart, "graph", use_graph, flush=True)
    op, st = _stream_op(problem)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B, restart=restart, tol=1e-10, maxit=100, use_graph=use_graph,
                                              device_steps=True)
    assert np.all(flags == 0) and np.all(relres <= 1e-10), relres
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-8, q
    assert max(_true_res(problem, X)) <= 1e-10
    assert abs(rv[-1] - relres.max()) <= 1e-12
    # the host-driven solver: the same iterations up to rounding
    Xh, fh, rh, ith, rvh = op.block_gmres(B, restart=restart, tol=1e-10, maxit=100)
    assert it[1] == ith[1] and abs(it[0] - ith[0]) <= 1
    assert _rel(X, Xh) <= 1e-9


def test_device_steps_graph_replay_equals_eager(problem):
    """Graph replays of the whole block step give the eager device-driven
    results exactly (restart 3: several cycles, every step graph replayed)."""
    print("case device-driven block graphs vs eager", flush=True)
    op, st = _stream_op(problem)
    B = problem[-1]
    Xe, fe, re, ite, rve = op.block_gmres(B, restart=3, tol=1e-9, maxit=100, device_steps=True)
    Xg, fg, rg, itg, rvg = op.block_gmres(B, restart=3, tol=1e-9, maxit=100, use_graph=True, device_steps=True)
    assert ite[1] >= 3 and itg == ite
    assert np.array_equal(Xg, Xe)
    assert np.array_equal(rvg, rve)


def test_device_steps_two_pass_and_complex64(problem, reference):
    from marie_amd import lib
    print("case device-driven block cgs2 and mode 1", flush=True)
    op, st = _stream_op(problem, lib.PREC_F64 | lib.PREC_F32)
    B = problem[-1]
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-10, maxit=100, cgs_passes=2, use_graph=True,
                                              device_steps=True)
    assert np.all(flags == 0) and np.all(relres <= 1e-10)
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-8, q
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-8, maxit=100, mode=1, inner_tol=1e-5,
                                              use_graph=True, device_steps=True)
    assert np.all(flags == 0) and np.all(relres <= 1e-8)
    assert max(_true_res(problem, X)) <= 1e-8
    for q in range(B.shape[1]):
        assert _rel(X[:, q], reference[:, q]) <= 1e-6


def test_device_steps_dependent_and_single_columns(problem, reference):
    """A rank-deficient block (repeated and proportional right-hand sides:
    the shifted CholeskyQR path and its third pass) and a single column."""
    print("case device-driven block dependent columns", flush=True)
    op, st = _stream_op(problem)
    b = problem[-1][:, 0]
    B = np.stack([b, 2j * b, b], axis=1)
    X, flags, relres, it, rv = op.block_gmres(B, restart=10, tol=1e-10, maxit=100, use_graph=True,
                                              device_steps=True)
    assert np.all(flags == 0)
    assert _rel(X[:, 0], reference[:, 0]) <= 1e-8
    assert _rel(X[:, 1], 2j * reference[:, 0]) <= 1e-8
    assert _rel(X[:, 2], reference[:, 0]) <= 1e-8
    X, flags, relres, it, rv = op.block_gmres(problem[-1][:, 2], restart=20, tol=1e-10, maxit=100,
                                              device_steps=True)
    assert flags[0] == 0 and relres[0] <= 1e-10
    assert _rel(X[:, 0], reference[:, 2]) <= 1e-8


def test_device_steps_argument_errors(problem):
    from marie_amd import lib
    op, st = _stream_op(problem)
    B = problem[-1][:, :2].copy()
    B[:, 1] = 0
    with pytest.raises(lib.MarieError):
        op.block_gmres(B, device_steps=True)
    wide = np.tile(problem[-1], (1, 9))          # 36 columns: above the device-driven limit of 32
    with pytest.raises(lib.MarieError):
        op.block_gmres(wide, restart=2, device_steps=True)


def test_device_steps_batch_of_uneven_problems():
    """Three bodies / frequencies of uneven sizes, each handle on its own
    stream, driven from one host thread with graph-captured steps: every
    problem gets the solution of its own device-driven solve (the same
    kernels in the same order: identical), and every column's true residual
    by the oracle's product meets tol."""
    import torch
    from marie_amd import lib
    print("case device-driven block batch", flush=True)
    probs = [_problem(10, 6.25e-3, 298.2e6), _problem(14, 5e-3, 128e6, eps=40.0, sig=0.3),
             _problem(12, 6.25e-3, 200e6, eps=60.0, sig=0.8)]
    ops, streams, Bs, Xs, nrhs, refs = [], [], [], [], [], []
    for p in probs:
        op, st = _stream_op(p)
        B = p[-1]
        ops.append(op)
        streams.append(st)
        Bd = torch.from_numpy(np.ascontiguousarray(B.T)).to('cuda:0')
        Bs.append(Bd)
        Xs.append(torch.zeros_like(Bd))
        nrhs.append(B.shape[1])
        Xr, fl, rr, it, rv = op.block_gmres(B, restart=8, tol=1e-9, maxit=100, use_graph=True, device_steps=True)
        assert np.all(fl == 0)
        refs.append(Xr)
    torch.cuda.synchronize()
    flags, relres, iters = lib.block_gmres_batch(ops, [b.data_ptr() for b in Bs], [x.data_ptr() for x in Xs],
                                                 nrhs, restart=8, tol=1e-9, maxit=100, use_graph=True,
                                                 device_steps=True)
    for i, p in enumerate(probs):
        assert np.all(flags[i] == 0) and np.all(relres[i] <= 1e-9)
        X = Xs[i].cpu().numpy().T
        assert np.array_equal(X, refs[i])
        assert max(_true_res(p, X)) <= 1e-8
    with pytest.raises(lib.MarieError):
        lib.block_gmres_batch([ops[0], ops[0]], [Bs[0].data_ptr()] * 2, [Xs[0].data_ptr()] * 2, [nrhs[0]] * 2,
                              device_steps=True)
