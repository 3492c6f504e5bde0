"""The reference's default self-terms (DIRECTFN) on the device.

marie_vie_singular_terms_directfn against the reference's own DIRECTFN
values (fixture tests/golden/ref_singular.npz, computed from
directfn_ws_*_const.cpp through their MEX argument orders) at Np = 10, 20,
30 for four (dx, f) cases; getOPERATORS(r, f, 'N') with its default method
against the oracle's assembly fed the same values; and at the 96x96x112
shape, the difference between the DIRECTFN and the machine-precision
operators against the oracle's, which is confined to the near-offset blocks
(nearby_correct.m) and so can be formed at full size.
"""
import os

import numpy as np
import pytest

from oracle import operators as oop

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'ref_singular.npz'))
C0 = 299792458.0


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _case(dx, f):
    return int(np.argmin(np.abs(G['cases'][:, 0] - dx) / dx + np.abs(G['cases'][:, 1] - f) / f))


def test_directfn_terms_match_reference():
    from marie_amd import lib
    print("case DIRECTFN self-terms vs the reference's values", flush=True)
    for (dx, f), ref in zip(G['cases'], G['directfn']):
        ko = 2 * np.pi / C0 * f
        for n, want in zip(G['np'], ref):
            got = lib.singular_terms_directfn(dx, ko, int(n))
            err = np.max(np.abs(got - want) / np.abs(want))
            assert err <= 1e-12, (dx, f, n, err)


@pytest.mark.parametrize("shape,dx,f", [((8, 7, 6), 6.25e-3, 298.2e6), ((6, 9, 7), 2e-3, 128e6)])
def test_getoperators_default_is_the_reference_default(shape, dx, f):
    from marie_amd import vie
    print("case getOPERATORS default", shape, flush=True)
    L, M, N = shape
    xs = [(np.arange(s) - (s - 1) / 2) * dx for s in shape]
    r = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
    fN = vie.getOPERATORS(r, f, 'N')
    ref = oop.get_operators(L, M, N, dx, f, 'N', method=G['directfn'][_case(dx, f), 1])
    assert _rel(fN, ref) <= 1e-12
    own = vie.getOPERATORS(r, f, 'N', singular='DEMCEM')
    assert _rel(own, oop.get_operators(L, M, N, dx, f, 'N')) <= 1e-12


def test_directfn_operator_at_config2_shape():
    from marie_amd import lib
    L, M, N, dx, f = 96, 96, 112, 2e-3, 298e6
    print("case DIRECTFN operator at 96x96x112", flush=True)
    ko = 2 * np.pi / C0 * f
    dfn = G['directfn'][_case(dx, f), 1]
    own = np.array(lib.singular_terms(dx, ko))
    ops = []
    for sv in (dfn, None):
        op = lib.Operator(L, M, N)
        op.assemble_N(dx, ko, sv)
        ops.append(op.get_fN())
        op.close()
    # oracle: the two operators differ only where nearby_correct.m places the
    # self-terms (P == 0 entries of the 8 near offsets)
    Ta, Tb = oop.nearby_tables(*dfn), oop.nearby_tables(*own)
    Pj = oop.vector_projection()
    dG = np.zeros((L, M, N, 6), complex)
    for key in Ta:
        Ca, Pa = (np.asarray(v) for v in Ta[key])
        Cb, _ = (np.asarray(v) for v in Tb[key])
        dJ = np.where(Pa == 0, Ca - Cb, 0)
        dG[key] = np.einsum('ij,ijk->k', dJ, Pj) / (4 * np.pi)
    dfN = oop.fft_operator(oop.circulant_nop(dG), workers=16)
    got = ops[0] - ops[1]
    scale = np.linalg.norm(ops[1])
    print("  |fN_DIRECTFN - fN_own| / |fN| = %.2e; device vs oracle difference %.2e of |fN|"
          % (np.linalg.norm(got) / scale, np.linalg.norm(got - dfN) / scale), flush=True)
    assert np.linalg.norm(got - dfN) <= 1e-12 * scale
    assert np.linalg.norm(got) > 1e-6 * scale


@pytest.mark.parametrize("op", ["N", "K"])
def test_getoperators_toeplitz_and_circulant_types(op):
    """getOPERATORS' 'T' and 'C' outputs against the oracle's assembly and
    circulant embedding (the same Toeplitz generator the 'FFTofC' path
    transforms), with the reference's default self-terms for N."""
    from marie_amd import vie
    shape, dx, f = (7, 6, 5), 6.25e-3, 298.2e6
    print("case getOPERATORS types T and C,", op, shape, flush=True)
    L, M, N = shape
    xs = [(np.arange(s) - (s - 1) / 2) * dx for s in shape]
    r = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
    method = G['directfn'][_case(dx, f), 1] if op == 'N' else 'own'
    for t, want_shape in (('T', (L, M, N)), ('C', (2 * L, 2 * M, 2 * N))):
        got = vie.getOPERATORS(r, f, op, type=t)
        ref = oop.get_operators(L, M, N, dx, f, op, type_=t, method=method)
        assert got.shape == want_shape + (6 if op == 'N' else 3,)
        assert _rel(got, ref) <= 1e-12, (op, t, _rel(got, ref))
    # the FFT of the 'C' output is the 'FFTofC' operator
    fC = np.fft.fftn(vie.getOPERATORS(r, f, op, type='C'), axes=(0, 1, 2))
    assert _rel(fC, vie.getOPERATORS(r, f, op)) <= 1e-13
