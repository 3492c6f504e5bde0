"""The multi-right-hand-side caller (fVIEsolver_Assembly + jvie2solver): the
bounding-box cut-out gives the same currents as the full-grid solve (the
operator restricted to the body's voxels depends only on voxel offsets), for
one port at a time (BiCGStab) and for all ports at once (block GMRES)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fviesolver_crop_and_ports():
    from marie_amd import coil, geometry, vie
    print("case fVIEsolver_Assembly", flush=True)
    n, dx, f = 22, 5e-3, 128e6
    r = geometry.cell_grid(n, n, n, dx, center=(0.004, -0.003, 0.002))
    em = vie.EMconstants(f)
    R = np.linalg.norm(r - np.array([0.01, 0.0, -0.005]), axis=-1)
    body = R < 0.03
    rng = np.random.default_rng(11)
    eps = np.where(body, 40 + 20 * rng.random(body.shape), 1.0)
    sig = np.where(body, 0.3 + 0.5 * rng.random(body.shape), 0.0)
    ports = coil.cylinder_loop_array(3, 0.08, 0.06, 0.03, seg_per_side=4)
    Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    RHBM = dict(r=r, epsilon_r=eps, sigma_e=sig)
    fVIEsolver, Gram, tau3, Scoord, dev = vie.fVIEsolver_Assembly(RHBM, f)
    # the cut-out box is the body's extent plus one voxel
    assert (dev.L, dev.M, dev.N) != (n, n, n) and dev.L * dev.M * dev.N < 0.7 * n ** 3
    idxS = np.nonzero(body.ravel(order='F'))[0]
    assert np.allclose(Scoord, r.reshape(-1, 3, order='F')[idxS], rtol=0, atol=1e-15)
    nS = idxS.size
    assert tau3.size == 3 * nS
    e_r = eps - 1j * sig / (em['eo'] * em['omega'])
    fN = vie.getOPERATORS(r, f, 'N')
    idxS3 = np.concatenate([idxS, n ** 3 + idxS, 2 * n ** 3 + idxS])
    V = np.stack([Gram * tau3 * E.ravel(order='F')[idxS3] for E in Es], axis=1)
    Xb = fVIEsolver(V, 1e-9)
    for p, E in enumerate(Es):
        Jref, flag, *_ = vie.JVIE_Solver(E, e_r, r, f, fN, 1e-11, 2, 'G')
        assert flag == 0
        ref = Jref.ravel(order='F')[idxS3]
        x = fVIEsolver(V[:, p], 1e-9, np.zeros(V.shape[0]))
        assert _rel(x, ref) <= 1e-6
        assert _rel(Xb[:, p], ref) <= 1e-6
    # a zero right-hand side returns zero (jvie2solver.m:30-33), also inside a batch
    assert np.all(fVIEsolver(np.zeros(V.shape[0]), 1e-6) == 0)
    V2 = V.copy()
    V2[:, 1] = 0
    X2 = fVIEsolver(V2, 1e-9)
    assert np.all(X2[:, 1] == 0) and _rel(X2[:, 0], Xb[:, 0]) <= 1e-6
    # an fN made for the full grid at this frequency is used when the box is large
    RHBM2 = dict(r=r, epsilon_r=np.where(np.ones_like(body), eps.max(), 1.0), sigma_e=np.full(body.shape, 0.4),
                 fN=fN, freqfN=f)
    s2, G2, t2, S2, dev2 = vie.fVIEsolver_Assembly(RHBM2, f)
    assert (dev2.L, dev2.M, dev2.N) == (n, n, n)
