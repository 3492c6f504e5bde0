"""Oracle: incident fields of current elements (eval_DGF.m:37-106).

Test infrastructure only (see oracle/__init__.py). Pinned by the static
(Biot-Savart) limit of H, Faraday's law curl E = -j w mu H and div E = 0
away from the sources (tests/test_oracle_excitation.py).
"""
import numpy as np

MU0 = 4 * np.pi * 1e-7
C0 = 299792458.0


def eval_dgf(J, Icoord, Ocoord, ko):
    """E, H (No x 3) of current elements J (Ni x 3) at Icoord, observed at Ocoord."""
    J = np.asarray(J, complex).reshape(-1, 3)
    Ic = np.asarray(Icoord, float).reshape(-1, 3)
    Oc = np.asarray(Ocoord, float).reshape(-1, 3)
    omega = ko * C0
    E = np.zeros(Oc.shape, complex)
    H = np.zeros(Oc.shape, complex)
    for jj in range(Ic.shape[0]):
        d = Oc - Ic[jj]
        R2 = np.sum(d * d, axis=1)
        R = np.sqrt(R2)
        R3 = R * R2
        chi = 1j * omega * MU0 * np.exp(-1j * ko * R) / (4 * np.pi * ko * ko * R3)
        Pp = 1j * ko * R + 1
        Q = ko * ko * R2 - Pp
        P = (Q - 2 * Pp) / R2
        Q = chi * Q
        P = chi * P
        E += (P[:, None] * d) * (d @ J[jj])[:, None] - Q[:, None] * J[jj][None, :]
        mult = np.exp(-1j * ko * R) * (1j * ko * R + 1) / (4 * np.pi * R3)
        H += mult[:, None] * np.cross(J[jj][None, :], d)
    return E, H
