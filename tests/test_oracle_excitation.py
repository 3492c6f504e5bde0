"""Pins of oracle.excitation.eval_dgf (the fields of eval_DGF.m)."""
import numpy as np

from oracle import excitation as ex


def test_static_limit_is_biot_savart():
    """k -> 0: H of a current element J at distance R is J x R / (4 pi R^3)."""
    rng = np.random.default_rng(1)
    J = rng.standard_normal((1, 3))
    src = np.zeros((1, 3))
    obs = rng.standard_normal((20, 3)) * 0.1 + 0.3
    E, H = ex.eval_dgf(J, src, obs, 1e-6)
    d = obs - src
    R = np.linalg.norm(d, axis=1)
    ref = np.cross(J[0][None, :], d) / (4 * np.pi * R[:, None] ** 3)
    assert np.max(np.abs(H - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_faraday_and_divergence():
    """curl E = -j w mu0 H and div E = 0 away from the element (central differences)."""
    rng = np.random.default_rng(2)
    ko = 2 * np.pi * 298e6 / ex.C0
    J = (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))) * 1e-2
    src = rng.standard_normal((3, 3)) * 0.02
    obs = rng.standard_normal((10, 3)) * 0.05 + np.array([0.2, 0.1, -0.15])
    h = 1e-6
    E, H = ex.eval_dgf(J, src, obs, ko)
    dE = []
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        dE.append((ex.eval_dgf(J, src, obs + e, ko)[0] - ex.eval_dgf(J, src, obs - e, ko)[0]) / (2 * h))
    curl = np.stack([dE[1][:, 2] - dE[2][:, 1], dE[2][:, 0] - dE[0][:, 2], dE[0][:, 1] - dE[1][:, 0]], axis=1)
    omega = ko * ex.C0
    assert np.max(np.abs(curl - (-1j * omega * ex.MU0) * H)) <= 1e-6 * np.max(np.abs(omega * ex.MU0 * H))
    div = dE[0][:, 0] + dE[1][:, 1] + dE[2][:, 2]
    assert np.max(np.abs(div)) <= 1e-6 * np.max(np.abs(dE[0]))
