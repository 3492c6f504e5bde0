"""CPU pinning of the oracle's operator assembly pieces.

Fixtures: tests/golden/ref_cubature.npz (written by tests/golden/make_golden.py
from the reference's own cubature_nop.cpp, DIRECTFN and DEMCEM sources and the
gauss_1d.m tables). Closed forms: the static coincident-square integral, the
projection identity, the circulant parity structure.
"""
import os

import numpy as np
import pytest

from oracle import mvp, singular
from oracle import operators as oop

GOLD = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'ref_cubature.npz'))
C0 = 299792458.0


@pytest.mark.parametrize("order", [4, 10, 20])
def test_gauss_1d_matches_reference_tables_bitwise(order):
    w, x = oop.gauss_1d(order)
    assert np.array_equal(x, GOLD['gauss_x_%d' % order])
    assert np.array_equal(w, GOLD['gauss_w_%d' % order])


@pytest.mark.parametrize("order", [1, 3, 7, 24])
def test_gauss_1d_polynomial_exactness(order):
    w, x = oop.gauss_1d(order)
    for deg in range(2 * order):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.sum(w * x ** deg) - exact) <= 1e-14


def test_gauss_2d_layout():
    Np, w, u, v = oop.gauss_2d(3)
    w1, x1 = oop.gauss_1d(3)
    assert Np == 9
    assert np.allclose(u[:3], (x1[0] + 1) / 2) and np.allclose(v[:3], (x1 + 1) / 2)   # u outer, v inner
    assert abs(w.sum() - 1.0) < 1e-15


def test_cubature_matches_reference_cpp():
    blocks, meta = GOLD['cub_blocks'], GOLD['cub_meta']
    worst = 0.0
    for B, (dx, f, order, mx, my, mz) in zip(blocks, meta):
        ko = 2 * np.pi / C0 * f
        got = oop.cubature_nop_batch(int(order), (np.array([mx, my, mz]) * dx)[None], oop.cube_faces(dx), ko,
                                     dx)[0]
        ok = np.isfinite(B)
        scale = np.max(np.abs(B[ok]))
        # entries the reference leaves infinite/NaN (coincident nodes) are the singular
        # ones that nearby_correct replaces; they are not compared
        worst = max(worst, np.max(np.abs(got[ok] - B[ok])) / scale)
    # summation order differs (numpy pairwise vs the C++ running sum over up to 160 000 terms)
    assert worst <= 2e-13


def test_vector_projection_closed_form():
    P = oop.vector_projection()
    e = np.eye(3)
    normals = [-e[0], e[0], -e[1], e[1], -e[2], e[2]]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k in range(6):
        for l in range(6):
            for c, (a, b) in enumerate(pairs):
                # (n_k x e_a).(n_l x e_b) = (n_k.n_l)(e_a.e_b) - (n_k.e_b)(e_a.n_l)
                val = np.dot(normals[k], normals[l]) * (a == b) - normals[k][b] * normals[l][a]
                assert P[k, l, c] == val
    assert set(np.unique(P)) <= {-1.0, 0.0, 1.0}


def test_circulant_structure():
    rng = np.random.default_rng(0)
    L, M, N = 4, 5, 6
    G = rng.standard_normal((L, M, N, 6)) + 1j * rng.standard_normal((L, M, N, 6))
    C = oop.circulant_nop(G)
    assert C.shape == (8, 10, 12, 6)
    assert np.all(C[L] == 0) and np.all(C[:, M] == 0) and np.all(C[:, :, N] == 0)
    for c in range(6):
        sx, sy, sz = oop.GP_COEFF['L'][c], oop.GP_COEFF['M'][c], oop.GP_COEFF['N'][c]
        i, j, k = 2, 3, 1
        assert C[2 * L - i, j, k, c] == sx * G[i, j, k, c]
        assert C[i, 2 * M - j, k, c] == sy * G[i, j, k, c]
        assert C[2 * L - i, 2 * M - j, 2 * N - k, c] == G[i, j, k, c]
        # product rule of the reflected blocks (gperiodic_coeff_nop.m)
        assert oop.GP_COEFF['LM'][c] == sx * sy and oop.GP_COEFF['LN'][c] == sx * sz
        assert oop.GP_COEFF['MN'][c] == sy * sz


def test_singular_static_closed_form_and_convergence():
    a = 1.5e-3
    assert abs(singular.st(0.0, a) - singular.st_static_closed_form(a)) <= 1e-15 * abs(
        singular.st_static_closed_form(a))
    ko = 2 * np.pi / C0 * 298e6
    t24 = np.array(singular.all_terms(ko, a, 24))
    t40 = np.array(singular.all_terms(ko, a, 40))
    assert np.max(np.abs(t24 - t40) / np.abs(t40)) <= 1e-14


def test_singular_against_demcem_and_directfn():
    """Own quadrature agrees with DEMCEM to DEMCEM's accuracy; DIRECTFN's edge-adjacent
    routine (called with the vertex order of directfn_ws_ea_const_mex.cpp:50) is not
    converged at Np = 20 and differs by ~3e-5 on the three entries that use it."""
    for (dx, f), dfn, dem in zip(GOLD['sing_cases'], GOLD['sing_directfn'], GOLD['sing_demcem']):
        ko = 2 * np.pi / C0 * f
        own = np.array(singular.all_terms(ko, dx))
        rel_dem = np.abs(own - dem) / np.abs(dem)
        assert np.all(rel_dem <= 1e-9), rel_dem
        rel_dfn = np.abs(own - dfn) / np.abs(dfn)
        assert np.all(rel_dfn[[0, 1, 2]] > 1e-6) and np.all(rel_dfn[[0, 1, 2]] < 1e-4)
        assert np.all(rel_dfn[[3, 4]] < 1e-6)


def test_assembly_far_entries_match_reference_cubature():
    """Far entries of G are the projected 4x4 cubature blocks of the reference C++."""
    blocks, meta = GOLD['cub_blocks'], GOLD['cub_meta']
    P = oop.vector_projection()
    L = M = N = 12
    for B, (dx, f, order, mx, my, mz) in zip(blocks, meta):
        if order != 4 or max(mx, my, mz) < 5 or dx != 6.25e-3:
            continue
        ko = 2 * np.pi / C0 * f
        G = oop.assembly_nop(L, M, N, ko, dx)
        g_ref = np.einsum('ij,ijk->k', B, P) / (4 * np.pi)
        g = G[int(mx), int(my), int(mz)]
        # the projection sums +-1 multiples of the face-pair entries, which cancel
        # strongly in the far field: compare on the scale of the entries themselves
        assert np.max(np.abs(g - g_ref)) <= 1e-13 * np.max(np.abs(B)) / (4 * np.pi)


def test_assembly_near_uses_singular_tables():
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    terms = singular.all_terms(ko, dx)
    T = oop.nearby_tables(*terms)
    C, Pm = T[(0, 0, 0)]
    assert C[0, 0] == terms[0] and Pm[0, 0] == 0
    assert C[0, 2] == terms[2]
    # each of the 8 near offsets replaces exactly the entries whose mask is 0
    for key, (Cm, Pm) in T.items():
        J = np.full((6, 6), 7.0 + 0j)
        out = oop.nearby_correct(key, J, T)
        assert np.all(out[Pm == 1] == 7.0) and np.all(out[Pm == 0] == Cm[Pm == 0])


def test_operator_is_complex_symmetric():
    """N built by the oracle assembly satisfies u^T N v = v^T N u (reciprocity)."""
    L = M = N = 6
    dx, f = 6.25e-3, 298.2e6
    fN = oop.get_operators(L, M, N, dx, f)
    rng = np.random.default_rng(1)
    idx = np.arange(3 * L * M * N)
    u = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    v = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    a = u @ mvp.mv_N(v, fN, L, M, N, idx)
    b = v @ mvp.mv_N(u, fN, L, M, N, idx)
    assert abs(a - b) <= 1e-12 * abs(a)
