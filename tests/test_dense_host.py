"""The library's host-side dense complex algebra (vie_dense.hpp: least
squares, thin QR, Cholesky solves, general eigenproblem), built with g++ and
checked against numpy on random and structured matrices."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def dh(tmp_path_factory):
    so = str(tmp_path_factory.mktemp('dense') / 'libdense_harness.so')
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-shared', '-fPIC', '-o', so,
                           os.path.join(HERE, 'dense_harness.cpp')])
    return ctypes.CDLL(so)


def _c(a):
    a = np.asfortranarray(np.asarray(a, np.complex128))
    return a, a.ctypes.data_as(ctypes.c_void_p)


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_lstsq_and_qr(dh):
    rng = np.random.default_rng(0)
    for m, n in ((7, 4), (51, 50), (30, 30), (120, 35)):
        A, pa = _c(_rand(rng, m, n))
        b, pb = _c(_rand(rng, m))
        x = np.zeros(n, np.complex128)
        assert dh.dh_lstsq(pa, m, n, pb, x.ctypes.data_as(ctypes.c_void_p)) == 0
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.linalg.norm(x - ref) <= 1e-12 * np.linalg.norm(ref)
        Q = np.zeros((m, n), np.complex128, order='F')
        R = np.zeros((n, n), np.complex128, order='F')
        dh.dh_qr_thin(pa, m, n, Q.ctypes.data_as(ctypes.c_void_p), R.ctypes.data_as(ctypes.c_void_p))
        assert np.allclose(Q @ R, A, atol=1e-12 * np.abs(A).max())
        assert np.allclose(Q.conj().T @ Q, np.eye(n), atol=1e-13)
        assert np.allclose(np.tril(R, -1), 0)
        X, px = _c(_rand(rng, 9, n))
        X0 = X.copy()
        assert dh.dh_right_solve_upper(px, 9, n, R.ctypes.data_as(ctypes.c_void_p)) == 0
        assert np.allclose(X @ R, X0, atol=1e-10 * np.abs(X0).max())


def test_chol_solve(dh):
    rng = np.random.default_rng(1)
    n, k = 40, 6
    G = _rand(rng, n + 3, n)
    B, pb = _c(G.conj().T @ G)
    A, pa = _c(_rand(rng, n, k))
    ref = np.linalg.solve(B, A)
    assert dh.dh_chol_solve(pb, n, pa, k) == 0
    assert np.linalg.norm(A - ref) <= 1e-11 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind", ['random', 'hessenberg', 'clustered', 'normal', 'jordan-ish', 'diag'])
def test_eig(dh, kind):
    rng = np.random.default_rng(hash(kind) % 2 ** 32)
    n = 60
    if kind == 'random':
        A = _rand(rng, n, n)
    elif kind == 'hessenberg':
        A = np.triu(_rand(rng, n, n), -1)
    elif kind == 'clustered':
        Q = np.linalg.qr(_rand(rng, n, n))[0]
        ev = np.concatenate([1e-3 * (1 + 0.01 * _rand(rng, 5)), 1 + 0.1 * _rand(rng, n - 5)])
        A = Q @ np.diag(ev) @ np.linalg.inv(Q + 0.1 * _rand(rng, n, n))
    elif kind == 'normal':
        Q = np.linalg.qr(_rand(rng, n, n))[0]
        A = Q @ np.diag(_rand(rng, n)) @ Q.conj().T
    elif kind == 'jordan-ish':
        A = np.diag(np.full(n, 2.0 + 1j)) + np.diag(np.full(n - 1, 1e-3), 1) + 1e-8 * _rand(rng, n, n)
    else:
        A = np.diag(_rand(rng, n))
    A, pa = _c(A)
    vals = np.zeros(n, np.complex128)
    vecs = np.zeros((n, n), np.complex128, order='F')
    assert dh.dh_eig(pa, n, vals.ctypes.data_as(ctypes.c_void_p), vecs.ctypes.data_as(ctypes.c_void_p)) == 0
    ref = np.linalg.eigvals(A)
    # every eigenvalue matches one of numpy's (sensitivity-scaled tolerance)
    tol = 1e-9 * np.abs(A).max() if kind != 'jordan-ish' else 1e-3
    for v in vals:
        assert np.min(np.abs(ref - v)) <= tol
    if kind != 'jordan-ish':
        res = np.linalg.norm(A @ vecs - vecs * vals, axis=0)
        assert np.max(res) <= 1e-10 * np.abs(A).max() * n
        assert np.allclose(np.linalg.norm(vecs, axis=0), 1)
