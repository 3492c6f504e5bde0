"""GPU parity of operator assembly (getOPERATORS 'N') and of the GMRES solve.

fN from the device assembly is compared with oracle.operators.get_operators
on small grids (<= 1e-12 relative, 2-norm over all six components); the
self-terms with oracle.singular (<= 1e-13); GMRES with oracle.krylov.pgmres
(same resvec to 1e-9, same solution to 1e-9, same flag / iteration counts).
"""
import numpy as np
import pytest

from oracle import fields, krylov, mvp, singular
from oracle import operators as oop

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dx,f", [(6.25e-3, 298.2e6), (2e-3, 128e6), (1.5e-3, 298e6)])
def test_singular_terms(dx, f):
    from marie_amd import lib
    print("case singular", dx, f, flush=True)
    ko = 2 * np.pi / C0 * f
    got = lib.singular_terms(dx, ko)
    ref = np.array(singular.all_terms(ko, dx))
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-13


@pytest.mark.parametrize("shape,dx,f", [((8, 7, 6), 6.25e-3, 298.2e6), ((6, 6, 9), 2e-3, 128e6),
                                        ((12, 12, 12), 5e-3, 64e6)])
def test_assembly_parity(shape, dx, f):
    from marie_amd import lib
    L, M, N = shape
    print("case assembly", shape, dx, f, flush=True)
    ko = 2 * np.pi / C0 * f
    op = lib.Operator(L, M, N)
    op.assemble_N(dx, ko)
    got = op.get_fN()
    ref = oop.get_operators(L, M, N, dx, f, 'N')
    assert _rel(got, ref) <= 1e-12
    # caller-supplied self-terms (the reference's DIRECTFN/DEMCEM values) are used verbatim
    sv = np.array(singular.all_terms(ko, dx)) * (1 + 1e-3)
    op.assemble_N(dx, ko, sv)
    got2 = op.get_fN()
    G = oop.assembly_nop(L, M, N, ko, dx, method='given', ref=_Given(sv))
    ref2 = oop.fft_operator(oop.circulant_nop(G))
    assert _rel(got2, ref2) <= 1e-12


@pytest.mark.parametrize("P", [(15, 13, 12), (16, 18, 11)])
def test_assembly_other_circulant(P):
    """Sides other than 2L (>= 2L-1): same embedding rule, same matvec as 2L."""
    from marie_amd import lib
    L, M, N = 8, 7, 6
    dx, f = 6.25e-3, 298.2e6
    print("case assembly circulant", P, flush=True)
    ko = 2 * np.pi / C0 * f
    op = lib.Operator(L, M, N, P)
    op.assemble_N(dx, ko)
    G = oop.assembly_nop(L, M, N, ko, dx)
    ref = oop.fft_operator(oop.circulant_nop(G, P))
    assert _rel(op.get_fN(), ref) <= 1e-12
    op2 = lib.Operator(L, M, N)
    op2.assemble_N(dx, ko)
    rng = np.random.default_rng(3)
    e_r = 1 + rng.random((L, M, N)) * 10 - 1j * rng.random((L, M, N))
    idx = np.arange(3 * L * M * N)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    ys = []
    for o in (op, op2):
        o.set_material(e_r, e_r - 1, dx ** 3)
        o.set_index(idx)
        ys.append(o.apply(x))
    assert _rel(ys[0], ys[1]) <= 1e-12


class _Given:
    def __init__(self, sv):
        self.sv = sv

    def singular_terms(self, ko, dx, method):
        return tuple(self.sv)


def _sphere_problem(n=12, dx=6.25e-3, f=298.2e6, eps=50.0, sig=0.5):
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    inside = X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2
    em = fields.em_constants(f)
    e_r = np.where(inside, eps - 1j * sig / (em['eo'] * em['omega']), 1.0 + 0j)
    Einc, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    return r, e_r, Einc, em


@pytest.mark.parametrize("restart", [50, 7])
def test_gmres_parity(restart):
    from marie_amd import lib
    n, dx, f = 12, 6.25e-3, 298.2e6
    print("case gmres restart", restart, flush=True)
    r, e_r, Einc, em = _sphere_problem(n, dx, f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    xr, flr, rr, itr, rvr = krylov.pgmres(A, b, restart, 1e-6, 200)
    op = lib.Operator(n, n, n)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    x, fl, rel, it, rv = op.gmres(b, restart=restart, tol=1e-6, maxit=200)
    assert fl == flr == 0
    assert it == itr
    assert rv.size == rvr.size
    assert np.max(np.abs(rv - rvr)) <= 1e-9
    assert _rel(x, xr) <= 1e-9
    assert abs(rel - rr) <= 1e-9


def test_gmres_mixed_precision():
    """complex64 matvecs + FP64 refinement reach the FP64 solution to the tolerance."""
    from marie_amd import lib
    n, dx, f = 14, 6.25e-3, 298.2e6
    print("case gmres mixed", flush=True)
    r, e_r, Einc, em = _sphere_problem(n, dx, f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    x64, fl64, rel64, _, _ = op.gmres(b, restart=50, tol=1e-8, maxit=200, mode=0)
    x32, fl32, rel32, it32, rv32 = op.gmres(b, restart=50, tol=1e-6, maxit=50, mode=1, inner_tol=1e-4)
    assert fl64 == 0 and fl32 == 0
    assert rel32 <= 1e-6
    assert _rel(x32, x64) <= 1e-5


def test_gmres_zero_rhs_and_maxit_flag():
    from marie_amd import lib
    print("case gmres zero rhs / flag", flush=True)
    n, dx, f = 8, 6.25e-3, 298.2e6
    r, e_r, Einc, em = _sphere_problem(n, dx, f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    op = lib.Operator(n, n, n)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    x, fl, rel, it, rv = op.gmres(np.zeros_like(b))
    assert fl == 0 and np.all(x == 0)
    x, fl, rel, it, rv = op.gmres(b, restart=2, tol=1e-12, maxit=1)
    assert fl == 1 and it[1] == 1


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_gmres_refined_modes_reach_tolerance(mode):
    """complex64 matvecs (mode 1: FP64 basis, mode 2: complex64 basis, mode 3:
    s-step cycles with a complex64 basis) with FP64 refinement: true residual
    <= tol by the oracle's FP64 product, and the solution within the FP64
    solve's own error of the oracle's pgmres solution converged to 1e-11."""
    from marie_amd import lib
    n, dx, f = 16, 6.25e-3, 298.2e6
    print("case gmres refined mode", mode, flush=True)
    r, e_r, Einc, em = _sphere_problem(n, dx, f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    xref, fl, rel, _, _ = krylov.pgmres(A, b, 50, 1e-11, 200)
    x64, fl64, rel64, _, _ = op.gmres(b, restart=50, tol=1e-6, maxit=200, mode=0)
    xm, flm, relm, itm, rvm = op.gmres(b, restart=50, tol=1e-6, maxit=50, mode=mode, inner_tol=1e-7)
    assert fl == 0 and fl64 == 0 and flm == 0
    assert relm <= 1e-6
    assert _rel(A(xm), b) <= 1e-6
    e64 = _rel(x64, xref)
    em_ = _rel(xm, xref)
    assert em_ <= max(5 * e64, 1e-6), (em_, e64)
    # resvec ends with the refined FP64 true residual
    assert abs(rvm[-1] - relm) <= 1e-9


@pytest.mark.parametrize("restart", [3, 8, 13, 40])
@pytest.mark.parametrize("adjoint", [False, True])
def test_gmres_sstep_restarts_and_adjoint(restart, adjoint):
    """Mode 3 (blocks of up to 8 complex64 products, block Gram-Schmidt and
    CholeskyQR on the device) with restarts below, at and between the block
    size, on A and on A^H: the true residual by the oracle's FP64 product
    (mv_AN 'notransp' / 'transp') is <= tol, the solution is within the FP64
    solve's own error of the oracle's pgmres solution, and resvec ends with the
    refined true residual."""
    import torch
    from marie_amd import lib
    n, dx, f = 12, 6.25e-3, 298.2e6
    print("case gmres s-step restart", restart, "adjoint", adjoint, flush=True)
    r, e_r, Einc, em = _sphere_problem(n, dx, f)
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    way = 'transp' if adjoint else 'notransp'
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, way, idx)
    xref, fl, rel, _, _ = krylov.pgmres(A, b, 50, 1e-11, 200)
    dev = torch.device('cuda:0')
    bd = torch.from_numpy(b).to(dev)
    out = {}
    for mode in (0, 3):
        xd = torch.zeros_like(bd)
        call = op.gmres_adjoint_device if adjoint else op.gmres_device
        out[mode] = call(bd.data_ptr(), xd.data_ptr(), restart=restart if mode else 50, tol=1e-6, maxit=400,
                         mode=mode, inner_tol=1e-7) + (xd.cpu().numpy(),)
    fl0, rel0, _, _, x0 = out[0]
    fl3, rel3, it3, rv3, x3 = out[3]
    assert fl == 0 and fl0 == 0 and fl3 == 0, (fl0, fl3, rel3)
    assert rel3 <= 1e-6 and _rel(A(x3), b) <= 1e-6
    assert _rel(x3, xref) <= max(5 * _rel(x0, xref), 1e-6)
    assert abs(rv3[-1] - rel3) <= 1e-9
    assert it3[0] <= restart


def test_gmres_sstep_odd_dof_count():
    """Mode 3 on a body with an odd number of DOFs (the two-element loads of
    the local Gram-Schmidt and the 16-byte column boundaries then leave one
    trailing element per column): true residual <= tol by the oracle's
    product, solution within the FP64 solve's error of the oracle's pgmres."""
    import torch
    from marie_amd import lib
    n, dx, f = 11, 6.25e-3, 298.2e6
    print("case gmres s-step odd DOF count", flush=True)
    r, e_r, Einc, em = _sphere_problem(n, dx, f)      # 485 voxels: 1455 DOFs
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=2)
    assert idx.size % 2 == 1
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    xref, fl, rel, _, _ = krylov.pgmres(A, b, 50, 1e-11, 200)
    x0, fl0, rel0, _, _ = op.gmres(b, restart=50, tol=1e-6, maxit=200, mode=0)
    for restart in (9, 40):
        x3, fl3, rel3, it3, rv3 = op.gmres(b, restart=restart, tol=1e-6, maxit=200, mode=3, inner_tol=1e-7)
        assert fl3 == 0 and rel3 <= 1e-6 and _rel(A(x3), b) <= 1e-6, restart
        assert _rel(x3, xref) <= max(5 * _rel(x0, xref), 1e-6), restart


def test_default_circulant_for_large_prime_factor():
    """A side whose 2L has a prime factor above 13 (L = 17: 34 = 2 * 17) gets
    the next 7-smooth circulant side (36) by default; the product equals the
    oracle's mv_AN on the reference's 2L embedding."""
    from marie_amd import lib
    L, M, N = 17, 6, 5
    dx, f = 6.25e-3, 298.2e6
    print("case default circulant", (L, M, N), flush=True)
    ko = 2 * np.pi / C0 * f
    op = lib.Operator(L, M, N)
    assert op.P == (36, 12, 10)
    op.assemble_N(dx, ko)
    fN = oop.get_operators(L, M, N, dx, f, 'N')          # 34 x 12 x 10 x 6
    rng = np.random.default_rng(4)
    e_r = 1 + rng.random((L, M, N)) * 10 - 1j * rng.random((L, M, N))
    Mr, Mc = np.ones((L, M, N)), (e_r - 1) / e_r
    idx = np.sort(rng.choice(3 * L * M * N, 900, replace=False))
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    op.set_material(Mr, Mc, dx ** 3)
    op.set_index(idx)
    assert _rel(op.apply(x), mvp.mv_AN(x, fN, Mr, Mc, dx ** 3, 'notransp', idx)) <= 1e-12
