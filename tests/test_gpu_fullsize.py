"""The configurations' real shapes on the GPU, and the largest grid.

96x96x112 voxels at 2 mm (a 192x192x224 circulant: radices 2, 3 and 7, the
LDS-stage engine) at 128 and 298 MHz: one FP64 product against the oracle's
mv_AN with the same operator (the device-assembled fN, itself parity-checked
at small sizes), and a GMRES solve whose true residual is recomputed. 128^3
voxels (256^3 circulant, the register engine): the size-independent
identities of the operator -- complex symmetry u^T N v = v^T N u and the
adjoint identity <A^H u, v> = <u, A v> -- and linearity.
"""
import numpy as np
import pytest

from oracle import mvp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _phantom(L, M, N, dx, f, scale):
    from marie_amd import geometry, vie
    r, eps, sig, _ = geometry.head_phantom(L, M, N, dx, seed=0, scale=scale)
    em = vie.EMconstants(f)
    e_r = eps - 1j * sig / (em['eo'] * em['omega'])
    Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    return r, e_r, Einc, em


@pytest.mark.parametrize("f", [128e6, 298e6])
def test_config2_product_and_solve(f):
    from marie_amd import lib, vie
    L, M, N, dx = 96, 96, 112, 2e-3
    print("case config-2 shape", f, flush=True)
    r, e_r, Einc, em = _phantom(L, M, N, dx, f, 1.0)
    Mr, Mc, Gram, idx1, b = vie.jvie_system(Einc, e_r, r, f, form=2)
    op = lib.Operator(L, M, N)
    assert op.P == (192, 192, 224)
    op.assemble_N(dx, em['ko'])
    fN = op.get_fN()
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(idx1.size) + 1j * rng.standard_normal(idx1.size)
    y = op.apply(x)
    yr = mvp.mv_AN(x, fN, Mr, Mc, Gram, 'notransp', idx1 - 1, workers=16)
    assert _rel(y, yr) <= 1e-12
    xs, flag, relres, it, resvec = op.gmres(b, restart=50, tol=1e-6, maxit=100)
    assert flag == 0 and relres <= 1e-6
    assert _rel(op.apply(xs), b) <= 1e-6            # the true residual, recomputed


def test_config3_identities():
    from marie_amd import lib, vie
    L = M = N = 128
    dx, f = 1.5e-3, 298e6
    print("case config-3 identities", flush=True)
    r, e_r, Einc, em = _phantom(L, M, N, dx, f, 0.85)
    Mr, Mc, Gram, idx1, b = vie.jvie_system(Einc, e_r, r, f, form=2)
    op = lib.Operator(L, M, N)
    op.assemble_N(dx, em['ko'])
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(idx1.size) + 1j * rng.standard_normal(idx1.size)
    v = rng.standard_normal(idx1.size) + 1j * rng.standard_normal(idx1.size)
    Nu, Nv = op.apply(u, lib.OP_N), op.apply(v, lib.OP_N)
    assert abs(u @ Nv - v @ Nu) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Nv)
    Av, Ahu = op.apply(v), op.apply(u, lib.OP_AN, True)
    assert abs(np.vdot(Ahu, v) - np.vdot(u, Av)) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Av)
    w = op.apply(2.0 * u - 0.5j * v)
    assert _rel(w, 2.0 * op.apply(u) - 0.5j * Av) <= 1e-13


def test_largest_grid_identities():
    """256^3 voxels (a 512^3 circulant: the register engine on 512-point lines,
    multi-GB work buffers and a 1.6 GB octant):
    complex symmetry of N, the adjoint identity of A, linearity, and a solve
    whose true residual is recomputed."""
    from marie_amd import lib, vie
    n = 256
    dx, f = 0.75e-3, 298e6
    print("case largest grid", n, flush=True)
    r, e_r, Einc, em = _phantom(n, n, n, dx, f, 0.85)
    Mr, Mc, Gram, idx1, b = vie.jvie_system(Einc, e_r, r, f, form=2)
    del r, Einc
    op = lib.Operator(n, n, n)
    assert op.P == (512, 512, 512)
    op.assemble_N(dx, em['ko'])
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)
    rng = np.random.default_rng(9)
    u = rng.standard_normal(idx1.size) + 1j * rng.standard_normal(idx1.size)
    v = rng.standard_normal(idx1.size) + 1j * rng.standard_normal(idx1.size)
    Nu, Nv = op.apply(u, lib.OP_N), op.apply(v, lib.OP_N)
    assert abs(u @ Nv - v @ Nu) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Nv)
    Av, Ahu = op.apply(v), op.apply(u, lib.OP_AN, True)
    assert abs(np.vdot(Ahu, v) - np.vdot(u, Av)) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Av)
    w = op.apply(2.0 * u - 0.5j * v)
    assert _rel(w, 2.0 * op.apply(u) - 0.5j * Av) <= 1e-13
    xs, flag, relres, it, resvec = op.gmres(b, restart=50, tol=1e-3, maxit=20)
    assert flag == 0 and relres <= 1e-3
    assert _rel(op.apply(xs), b) <= 1e-3
