"""Pins of the oracle's K operator (assembly_kop.m, cubature_kop.cpp,
cubature_kop_near_sym.m, nearby_correct_kop.m, circulant_kop.m, mv_K.m)."""
import os

import numpy as np
import pytest

from oracle import mvp
from oracle import operators as oop

HERE = os.path.dirname(os.path.abspath(__file__))
C0 = 299792458.0


def test_cubature_kop_matches_reference_fixture():
    """Far/medium formula against the reference's own cubature_kop.cpp output."""
    d = np.load(os.path.join(HERE, 'golden', 'ref_cubature_kop.npz'))
    for blk, (dx, f, order, mx, my, mz) in zip(d['kop_blocks'], d['kop_meta']):
        ko = 2 * np.pi / C0 * f
        got = oop.cubature_kop_batch(int(order), np.array([mx, my, mz]) * dx, oop.cube_faces(dx), ko)[0]
        assert np.max(np.abs(got - blk)) <= 1e-12 * np.max(np.abs(blk)), (dx, f, order, mx, my, mz)


def test_near_formula_is_the_same_integral():
    """At edge and vertex contact both formulas apply: the near one (21 pairs,
    completion, coincident-panel zeroing, row combination) must agree with the
    fixture-pinned 12-pair formula at the same 20x20 rule."""
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    Rf = oop.cube_faces(dx)
    for m in [(1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]:
        near = oop.cubature_kop_near(20, m, Rf, ko, dx)
        far = oop.cubature_kop_batch(20, np.array(m, float) * dx, Rf, ko)[0]
        assert np.max(np.abs(near - far)) <= 1e-6 * np.max(np.abs(far)), m


def test_far_field_limit():
    """Far from the source, IK(m) = -dx^6 grad(exp(-jkR)/R)/(4 pi) times the
    mean-value factor of two cube averages of a Helmholtz field,
    1 - (k dx)^2/12, plus an O((dx/R)^2) remainder: fixes the sign, the
    1/(jk)^2 dx^4/(4 pi) normalisation and the component order."""
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    Rf = oop.cube_faces(dx)
    for scale in (1, 2, 4):
        m = np.array([15, 10, 5]) * scale
        rm = m * dx
        ik = oop.cubature_kop_batch(10, rm, Rf, ko)[0] * dx ** 4 / (1j * ko) ** 2 / (4 * np.pi)
        R = np.linalg.norm(rm)
        grad = -(1 + 1j * ko * R) * np.exp(-1j * ko * R) / R ** 3 * rm
        ref = -dx ** 6 * grad / (4 * np.pi) * (1 - (ko * dx) ** 2 / 12)
        e = np.max(np.abs(ik - ref)) / np.max(np.abs(ref))
        assert e <= 1e-3 * (dx / R) ** 2, (scale, e)


def test_assembly_symmetry_and_self_term():
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    G = oop.assembly_kop(6, 5, 5, ko, dx)
    scale = np.max(np.abs(G))
    assert np.max(np.abs(G[0, 0, 0])) <= 1e-12 * scale          # odd kernel over a cube with itself
    # on an axis only the axial component survives
    assert np.max(np.abs(G[1:, 0, 0, 1:])) <= 1e-12 * scale
    assert np.max(np.abs(G[0, 1:, 0, [0, 2]])) <= 1e-12 * scale
    assert np.max(np.abs(G[0, 0, 1:, :2])) <= 1e-12 * scale
    # the same face-adjacent value on each axis (cube symmetry)
    assert abs(G[1, 0, 0, 0] - G[0, 1, 0, 1]) <= 1e-12 * scale
    assert abs(G[1, 0, 0, 0] - G[0, 0, 1, 2]) <= 1e-12 * scale


@pytest.mark.parametrize("P", [None, (9, 11, 10)])
def test_mv_K_matches_dense_toeplitz(P):
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    L, M, N = 5, 6, 5
    G = oop.assembly_kop(L, M, N, ko, dx)
    fK = oop.fft_operator(oop.circulant_kop(G, P))
    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(3 * L * M * N, 200, replace=False))
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    y = mvp.mv_K(x, fK, L, M, N, idx)
    yd = mvp.toeplitz_apply_dense_k(x, G, idx)
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) <= 1e-12
