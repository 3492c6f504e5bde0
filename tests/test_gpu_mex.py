"""The MEX gateways (marie_amd/mex) called as MATLAB calls them, through the
MEX API emulation of tests/mex, against the oracle.

mv_AN_hip(JIn0, fN, Mr, Mc, Gram, transp_flag, idx) must return mv_AN.m's
column for both flags, keep its device operator across calls with the same
arrays and follow a changed material; jvie_gmres_hip must return pgmres'
[x, flag, relres, iter, resvec] for the JVIE system (pgmres.m:70-181);
mv_N_hip(JIn0, fN, r, idx), mv_L_hip(JIn0, fN, r, Gram, idx) and
mv_K_hip(JIn0, fK, r, idx) must return the columns of mv_N.m, mv_L.m and
mv_K.m; jvie_bicgstab_hip must follow bicgstab's recurrences and outputs
(JVIE_Solver.m:451-497); getOPERATORS_hip(r, f, op, type, singular) must
return getOPERATORS.m's arrays.
"""
import os
import sys

import numpy as np
import pytest

from oracle import fields, krylov, mvp
from oracle import operators as oop

sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'mex'))

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _problem(n=10, dx=6.25e-3, f=298.2e6, form=1):
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    em = fields.em_constants(f)
    e_r = np.where(X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2, 50 - 1j * 0.5 / (em['eo'] * em['omega']), 1.0)
    Einc, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=form)
    return fN, Mr, Mc, Gram, idx, b


def test_mv_AN_gateway():
    import emul
    print("case MEX gateway mv_AN_hip", flush=True)
    g = emul.Gateway('mv_AN_hip')
    fN, Mr, Mc, Gram, idx, b = _problem(form=1)
    idx1 = (idx + 1).astype(float)                       # MATLAB's 1-based indices
    rng = np.random.default_rng(1)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    for flag in ('notransp', 'transp'):
        y = g(x, fN, Mr, Mc, Gram, flag, idx1)
        assert y.shape == (idx.size, 1)
        assert _rel(y[:, 0], mvp.mv_AN(x, fN, Mr, Mc, Gram, flag, idx)) <= 1e-12, flag
    # the same arrays again (cached operator), then a new material
    y2 = g(x, fN, Mr, Mc, Gram, 'notransp', idx1)
    assert _rel(y2[:, 0], mvp.mv_AN(x, fN, Mr, Mc, Gram, 'notransp', idx)) <= 1e-12
    Mc2 = 0.5 * Mc
    y3 = g(x, fN, Mr, Mc2, Gram, 'notransp', idx1)
    assert _rel(y3[:, 0], mvp.mv_AN(x, fN, Mr, Mc2, Gram, 'notransp', idx)) <= 1e-12
    # MATLAB-style errors: arity, flag, sizes
    with pytest.raises(emul.MexError, match='MARIE:hip:nargin'):
        g(x, fN, Mr, Mc, Gram, 'notransp')
    with pytest.raises(emul.MexError, match='MARIE:hip:transp'):
        g(x, fN, Mr, Mc, Gram, 'sideways', idx1)
    with pytest.raises(emul.MexError, match='MARIE:hip:size'):
        g(x[:-1], fN, Mr, Mc, Gram, 'notransp', idx1)
    g('clear', nargout=0)


def test_jvie_gmres_gateway():
    import emul
    print("case MEX gateway jvie_gmres_hip", flush=True)
    g = emul.Gateway('jvie_gmres_hip')
    fN, Mr, Mc, Gram, idx, b = _problem(form=2)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    xr, flr, rr, itr, rvr = krylov.pgmres(A, b, 20, 1e-8, 20)
    x, flag, relres, it, resvec = g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), 20, 1e-8, 20, nargout=5)
    assert flag[0, 0] == flr == 0
    assert list(it.ravel().astype(int)) == list(itr)
    assert resvec.shape[1] == 1 and resvec.size == rvr.size
    assert np.max(np.abs(resvec[:, 0] - rvr)) <= 1e-9
    assert _rel(x[:, 0], xr) <= 1e-9 and abs(relres[0, 0] - rr) <= 1e-9
    # a start vector
    x2, flag2 = g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), 20, 1e-8, 20, xr, nargout=2)
    assert flag2[0, 0] == 0 and _rel(x2[:, 0], xr) <= 1e-7
    # the refined precision (complex64 products in block-orthogonalised
    # cycles, FP64 refinement): pgmres' flag/relres on the true residual
    x3, flag3, relres3, it3, rv3 = g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), 40, 1e-6, 20, np.zeros(0),
                                     'fp32-refined', nargout=5)
    assert flag3[0, 0] == 0 and relres3[0, 0] <= 1e-6
    assert _rel(A(x3[:, 0]), b) <= 1e-6 and abs(rv3[-1, 0] - relres3[0, 0]) <= 1e-9
    assert _rel(x3[:, 0], xr) <= 1e-4
    with pytest.raises(emul.MexError, match='MARIE:hip:precision'):
        g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), 40, 1e-6, 20, np.zeros(0), 'fp16')
    g('clear', nargout=0)


def test_mv_N_L_K_gateways():
    import emul
    print("case MEX gateways mv_N_hip, mv_L_hip, mv_K_hip", flush=True)
    n, dx, f = 9, 6.25e-3, 298.2e6
    x1 = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x1, x1, x1, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    em = fields.em_constants(f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    fK = oop.fft_operator(oop.circulant_kop(oop.assembly_kop(n, n, n, em['ko'], dx)))
    rng = np.random.default_rng(5)
    # E_field_Nop.m:69 / H_field_Kop.m:69 use every DOF; a subset checks the scatter too
    for idx in (np.arange(3 * n ** 3), np.sort(rng.choice(3 * n ** 3, 1500, replace=False))):
        idx1 = (idx + 1).astype(float)
        x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
        gN, gL, gK = emul.Gateway('mv_N_hip'), emul.Gateway('mv_L_hip'), emul.Gateway('mv_K_hip')
        y = gN(x, fN, r, idx1)
        assert y.shape == (idx.size, 1)
        assert _rel(y[:, 0], mvp.mv_N(x, fN, n, n, n, idx)) <= 1e-12
        for gram in (dx ** 3, 2 * dx ** 3):              # a changed Gram is applied
            y = gL(x, fN, r, gram, idx1)
            assert _rel(y[:, 0], mvp.mv_L(x, fN, n, n, n, gram, idx)) <= 1e-12
        y = gK(x, fK, r, idx1)
        assert _rel(y[:, 0], mvp.mv_K(x, fK, n, n, n, idx)) <= 1e-12
        with pytest.raises(emul.MexError, match='MARIE:hip:r'):
            gN(x, fN, r[..., 0], idx1)
        with pytest.raises(emul.MexError, match='MARIE:hip:fN'):
            gK(x, fN, r, idx1)
        with pytest.raises(emul.MexError, match='MARIE:hip:size'):
            gL(x[:-1], fN, r, dx ** 3, idx1)
        for g in (gN, gL, gK):
            g('clear', nargout=0)


@pytest.mark.parametrize("precond,form", [(0, 2), (1, 1)])
def test_jvie_bicgstab_gateway(precond, form):
    import emul
    print("case MEX gateway jvie_bicgstab_hip precond", precond, "form", form, flush=True)
    g = emul.Gateway('jvie_bicgstab_hip')
    fN, Mr, Mc, Gram, idx, b = _problem(form=form)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    M = (lambda v: mvp.mv_AN(v, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, 'notransp', idx)) if precond else None
    tol = 1e-9
    xr, flr, rr, itr, rvr = krylov.bicgstab(A, b, tol, 400, M=M)
    x, flag, relres, it, resvec = g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), tol, 400, precond, nargout=5)
    assert flag[0, 0] == flr == 0 and relres[0, 0] <= tol
    assert abs(it[0, 0] - itr) <= 4.0             # a short recurrence: counts may part by a few steps
    k = min(resvec.shape[0], rvr.size, 12)
    assert np.max(np.abs(resvec[:k, 0] - rvr[:k]) / rvr[0]) <= 1e-10
    assert _rel(x[:, 0], xr) <= 100 * tol
    assert _rel(A(x[:, 0]), b) <= tol * 1.01
    with pytest.raises(emul.MexError, match='MARIE:hip:nargin'):
        g(b, fN, Mr, Mc, Gram, (idx + 1).astype(float), tol)
    g('clear', nargout=0)


def test_getOPERATORS_gateway():
    import emul
    from marie_amd import vie
    print("case MEX gateway getOPERATORS_hip", flush=True)
    g = emul.Gateway('getOPERATORS_hip')
    n, dx, f = 7, 6.25e-3, 298.2e6
    x1 = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x1, x1, x1 + 0.01, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    # 'DEMCEM': the library's machine-precision self-terms, the oracle's 'own'
    for typ in ('FFTofC', 'T', 'C'):
        ref = oop.get_operators(n, n, n, dx, f, 'N', typ)
        out = g(r, f, 'N', typ, 'DEMCEM')
        assert out.shape == ref.shape, typ
        assert _rel(out, ref) <= 1e-12, typ
        refk = oop.get_operators(n, n, n, dx, f, 'K', typ)
        outk = g(r, f, 'K', typ)
        assert outk.shape == refk.shape and _rel(outk, refk) <= 1e-12, typ
    # the default ('N', 'FFTofC', 'DIRECTFN'): the reference's DIRECTFN self-terms
    out = g(r, f)
    assert _rel(out, vie.getOPERATORS(r, f)) <= 1e-14
    assert 1e-9 < _rel(out, oop.get_operators(n, n, n, dx, f, 'N')) < 1e-3
    with pytest.raises(emul.MexError, match='MARIE:hip:type'):
        g(r, f, 'N', 'X')
    with pytest.raises(emul.MexError, match='MARIE:hip:r'):
        g(r[..., 0], f)
