"""The K operator on the device: assembly and product parity with the oracle
(itself pinned by the reference's cubature_kop.cpp fixtures, the far-field
limit and a dense Toeplitz sum), and the H field of the dielectric sphere
against the curl of the Mie interior field."""
import numpy as np
import pytest

from oracle import fields, mie, mvp
from oracle import operators as oop

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape,P,dx,f", [((8, 7, 6), None, 6.25e-3, 298.2e6), ((6, 6, 9), None, 2e-3, 128e6),
                                          ((8, 7, 6), (15, 13, 12), 6.25e-3, 298.2e6)])
def test_kop_assembly_parity(shape, P, dx, f):
    from marie_amd import lib
    L, M, N = shape
    print("case K assembly", shape, P, flush=True)
    ko = 2 * np.pi / C0 * f
    op = lib.Operator(L, M, N, P)
    op.assemble_K(dx, ko)
    got = op.get_fK()
    ref = oop.fft_operator(oop.circulant_kop(oop.assembly_kop(L, M, N, ko, dx), P))
    assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize("shape,P", [((8, 7, 6), None), ((6, 5, 7), (11, 10, 13)), ((16, 16, 16), None)])
def test_mv_K_parity(shape, P):
    from marie_amd import lib
    L, M, N = shape
    print("case mv_K", shape, P, flush=True)
    dx, f = 6.25e-3, 298.2e6
    ko = 2 * np.pi / C0 * f
    fK = oop.fft_operator(oop.circulant_kop(oop.assembly_kop(L, M, N, ko, dx), P))
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(3 * L * M * N, (3 * L * M * N) // 2, replace=False))
    x = rng.standard_normal((idx.size, 2)) + 1j * rng.standard_normal((idx.size, 2))
    op = lib.Operator(L, M, N, P, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fK(fK)
    op.set_index(idx)
    y = op.apply(x, lib.OP_K)
    for q in range(2):
        yr = mvp.mv_K(x[:, q], fK, L, M, N, idx)
        assert _rel(y[:, q], yr) <= 1e-12
    # the stored operator round-trips (octant storage when the parity holds to
    # 1e-14 of the largest entry, the FFT's rounding asymmetry)
    assert _rel(op.get_fK(), fK) <= 1e-13
    import torch
    xd = torch.from_numpy(np.ascontiguousarray(x[:, 0])).to('cuda:0').to(torch.complex64)
    yd = torch.empty_like(xd)
    op.apply_device(xd.data_ptr(), yd.data_ptr(), lib.OP_K, False, lib.PREC_F32)
    torch.cuda.synchronize()
    assert _rel(yd.cpu().numpy().astype(complex), mvp.mv_K(x[:, 0], fK, L, M, N, idx)) <= 2e-6


def test_kop_errors():
    from marie_amd import lib
    op = lib.Operator(8, 8, 8)
    op.set_index(np.arange(10))
    with pytest.raises(lib.MarieError):
        op.apply(np.ones(10), lib.OP_K)          # no fK yet
    op.assemble_K(6.25e-3, 2 * np.pi / C0 * 298e6)
    with pytest.raises(lib.MarieError):
        op.apply(np.ones(10), lib.OP_K, transp=True)
    with pytest.raises(lib.MarieError):
        op.assemble_K(6.25e-3, 0.0)


def _mie_H(pts, a, eps, sig, f, h=1e-6):
    """H = curl E / (-j omega mu0) of the Mie interior field (exp(+j omega t)), by
    central differences of the closed-form series."""
    em = fields.em_constants(f)
    E = lambda p: mie.mie_interior_field(p, a, eps, sig, f)
    curl = np.zeros(pts.shape, complex)
    d = {}
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = h
        d[ax] = (E(pts + e) - E(pts - e)) / (2 * h)      # dE/d x_ax, all components
    curl[:, 0] = d[1][:, 2] - d[2][:, 1]
    curl[:, 1] = d[2][:, 0] - d[0][:, 2]
    curl[:, 2] = d[0][:, 1] - d[1][:, 0]
    return curl / (-1j * em['omega'] * 4 * np.pi * 1e-7)


def test_sphere_H_field_against_mie():
    """Lossless eps_r = 4 sphere (radius 0.1 m, 298 MHz): the VIE total H
    (Hinc + K J / Gram, H_field_Kop.m) in the inner voxels against the curl of
    the Mie field; the discretisation error decreases from 12.5 mm to 6.25 mm."""
    from marie_amd import geometry, vie
    f, a, eps, sig = 298e6, 0.1, 4.0, 0.0
    errs = []
    for n, dx in ((16, 12.5e-3), (32, 6.25e-3)):
        print("case sphere H", n, dx, flush=True)
        r = geometry.cell_grid(n, n, n, dx)
        em = vie.EMconstants(f)
        R = np.linalg.norm(r, axis=-1)
        e_r = np.where(R < a, eps - 1j * sig / (em['eo'] * em['omega']), 1.0 + 0j)
        Einc, Hinc = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
        fN = vie.getOPERATORS(r, f, 'N', return_operator=True)
        fK = vie.getOPERATORS(r, f, 'K', return_operator=True)
        J, flag, relres, *_ = vie.JVIE_Solver(Einc, e_r, r, f, fN, tol=1e-8, solver='G')
        assert flag == 0
        H = vie.H_field_Kop(J, fK, dx ** 3, Hinc)
        inner = R < 0.6 * a
        Hm = _mie_H(r[inner], a, eps, sig, f)
        errs.append(np.linalg.norm(H[inner] - Hm) / np.linalg.norm(Hm))
        print("  H error %.4f" % errs[-1], flush=True)
    assert errs[1] < errs[0] and errs[1] <= 0.1, errs


def test_scat_solver_outputs():
    """Scat_Solver's outputs on a small lossy sphere: B = mu0 H, Pabs = 0.5 sum
    sigma |E|^2 dV, Gsar = sum SAR dV."""
    from marie_amd import geometry, vie
    f = 298e6
    n, dx = 16, 8e-3
    r = geometry.cell_grid(n, n, n, dx)
    em = vie.EMconstants(f)
    R = np.linalg.norm(r, axis=-1)
    body = R < 0.055
    eps = np.where(body, 50.0, 1.0)
    sig = np.where(body, 0.5, 0.0)
    rho = np.where(body, 1000.0, 0.0)
    Einc, Hinc = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    J, S, E, B, Gsar, Pabs, fN, fK = vie.Scat_Solver(Einc, Hinc, f, r, eps, sig, rho, tol=1e-9)
    H = vie.H_field_Kop(J, fK, dx ** 3, Hinc)
    assert np.allclose(B, em['mu'] * H)
    ref = 0.5 * np.sum(sig[..., None] * np.abs(E) ** 2) * dx ** 3
    assert abs(Pabs - ref) <= 1e-6 * ref
    assert abs(Gsar - np.sum(S) * dx ** 3) <= 1e-12 * Gsar
    assert np.all(np.isfinite(vie.B1_plus(B)))
