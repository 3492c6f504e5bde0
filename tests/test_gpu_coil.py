"""Coil-port excitations on the device: the eval_DGF kernel against the
oracle (pinned by Biot-Savart and Faraday's law), CCL_Excitation of a
16-loop cylindrical array, and the 16 ports solved together by block GMRES."""
import numpy as np
import pytest

from oracle import excitation as oex

pytestmark = pytest.mark.gpu


def test_dgf_kernel_parity():
    from marie_amd import lib
    print("case dgf parity", flush=True)
    rng = np.random.default_rng(5)
    ko = 2 * np.pi * 298e6 / oex.C0
    starts = [0, 7, 7, 40]                      # three ports, one without segments
    sc = rng.standard_normal((40, 3)) * 0.05 + np.array([0.2, 0.0, 0.0])
    sj = (rng.standard_normal((40, 3)) + 1j * rng.standard_normal((40, 3))) * 1e-2
    pts = rng.standard_normal((1000, 3)) * 0.05
    E, H = lib.dgf_fields(sc, sj, starts, pts, ko, 2 * np.pi * 298e6 * oex.MU0)
    for p in range(3):
        a, b = starts[p], starts[p + 1]
        if a == b:
            assert np.all(E[p] == 0) and np.all(H[p] == 0)
            continue
        Er, Hr = oex.eval_dgf(sj[a:b], sc[a:b], pts, ko)
        assert np.max(np.abs(E[p] - Er)) <= 1e-12 * np.max(np.abs(Er))
        assert np.max(np.abs(H[p] - Hr)) <= 1e-12 * np.max(np.abs(Hr))


def test_ccl_excitation_loop_array():
    from marie_amd import coil, geometry, vie
    print("case ccl array", flush=True)
    f = 128e6
    em = vie.EMconstants(f)
    r = geometry.cell_grid(12, 12, 10, 8e-3)
    ports = coil.cylinder_loop_array(16, 0.15, 0.1, 0.05, seg_per_side=6)
    Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    pts = r.reshape(-1, 3, order='F')
    for p in (0, 5, 15):
        rows = ports[p]
        Er, Hr = oex.eval_dgf(rows[:, 0:3] - rows[:, 3:6], (rows[:, 0:3] + rows[:, 3:6]) / 2, pts, em['ko'])
        assert np.max(np.abs(Es[p].reshape(-1, 3, order='F') - Er)) <= 1e-12 * np.max(np.abs(Er))
        assert np.max(np.abs(Hs[p].reshape(-1, 3, order='F') - Hr)) <= 1e-12 * np.max(np.abs(Hr))
    # a closed loop is divergence free: the sum of its current elements vanishes
    for rows in ports:
        assert np.max(np.abs(np.sum(rows[:, 0:3] - rows[:, 3:6], axis=0))) <= 1e-15
    # rotating the array by one port rotates the field pattern (16-fold symmetry)
    rot = np.array([[np.cos(2 * np.pi / 16), -np.sin(2 * np.pi / 16), 0], [np.sin(2 * np.pi / 16),
                    np.cos(2 * np.pi / 16), 0], [0, 0, 1]])
    pt = np.array([[0.02, 0.01, 0.005]])
    E0, _ = oex.eval_dgf(ports[0][:, 0:3] - ports[0][:, 3:6], (ports[0][:, 0:3] + ports[0][:, 3:6]) / 2, pt,
                         em['ko'])
    E1, _ = oex.eval_dgf(ports[1][:, 0:3] - ports[1][:, 3:6], (ports[1][:, 0:3] + ports[1][:, 3:6]) / 2,
                         pt @ rot.T, em['ko'])
    assert np.max(np.abs(E1[0] - rot @ E0[0])) <= 1e-10 * np.max(np.abs(E0))


def test_coil_ports_block_solve():
    """16 coil ports of a small lossy body in one block GMRES; every port's true
    residual <= tol and each port within tol-level of its single solve."""
    from marie_amd import coil, geometry, lib, vie
    print("case coil block", flush=True)
    f = 128e6
    n, dx = 14, 8e-3
    r = geometry.cell_grid(n, n, n, dx)
    em = vie.EMconstants(f)
    R = np.linalg.norm(r, axis=-1)
    body = R < 0.05
    e_r = np.where(body, 60 - 1j * 0.6 / (em['eo'] * em['omega']), 1.0 + 0j)
    ports = coil.cylinder_loop_array(16, 0.15, 0.1, 0.05)
    Es, _ = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    cols = [vie.jvie_system(E, e_r, r, f, form=2) for E in Es]
    Mr, Mc, Gram, idx1, _ = cols[0]
    B = np.stack([c[4] for c in cols], axis=1)
    op = vie.getOPERATORS(r, f, 'N', return_operator=True)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)
    X, flags, relres, it, rv = op.block_gmres(B, restart=20, tol=1e-8, maxit=50)
    assert np.all(flags == 0) and np.all(relres <= 1e-8)
    for q in (0, 7, 15):
        x, fl, rel, _, _ = op.gmres(B[:, q], restart=50, tol=1e-10, maxit=100)
        assert np.linalg.norm(X[:, q] - x) / np.linalg.norm(x) <= 1e-6
