"""CPU pinning of the Mie interior-field oracle (config-1 physics reference).

The interior field (c_n, d_n), the scattered field (a_n, b_n) and the closed-form
incident wave must satisfy the interface conditions on r = a: continuous
tangential E and continuous normal D (m^2 E_1.n = (E_i + E_s).n). A wrong sign,
index or coefficient breaks them. The quasi-static limit of a small lossless
sphere (uniform 3/(eps+2)) pins the normalisation.
"""
import numpy as np
import pytest

from oracle import mie

MU0 = 4 * np.pi * 1e-7
C0 = 299792458.0
EO = 1.0 / C0 ** 2 / MU0


@pytest.mark.parametrize("eps,sig,f,a", [(50.0, 0.5, 298.2e6, 0.1), (52.0, 0.55, 298.2e6, 0.1),
                                         (4.0, 0.0, 1e9, 0.05), (80.0, 2.0, 128e6, 0.09)])
def test_interface_conditions(eps, sig, f, a):
    w = 2 * np.pi * f
    k = w / C0
    m = np.sqrt(np.conj(eps - 1j * sig / (EO * w)))
    rng = np.random.default_rng(0)
    v = rng.standard_normal((16, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    E1, Es, Ei = mie.bh_fields(a * v, a, m, k, 50)
    Eo = Es + Ei
    tang = lambda E: E - np.sum(E * v, -1)[:, None] * v
    scale = np.max(np.abs(Eo))
    assert np.max(np.abs(tang(E1) - tang(Eo))) <= 1e-12 * scale
    assert np.max(np.abs(m * m * np.sum(E1 * v, -1) - np.sum(Eo * v, -1))) <= 1e-12 * scale


def test_quasistatic_limit_and_convention():
    eps, a, f = 4.0, 1e-3, 1e6
    pts = np.array([[0.0, 0.0, 0.0], [2e-4, 1e-4, -3e-4], [-5e-4, 2e-4, 1e-4]])
    E = mie.mie_interior_field(pts, a, eps, 0.0, f)
    # first correction is the O(k r) magnetic-dipole term (k r ~ 1e-5 here)
    assert np.allclose(E[:, 0], 3.0 / (eps + 2.0), rtol=2e-5)
    assert np.allclose(E[:, 1:], 0.0, atol=2e-5)
    # MARIE's exp(+jwt) convention: conjugate of the B&H field
    w = 2 * np.pi * 298e6
    p = np.array([[0.01, 0.02, -0.03]])
    Em = mie.mie_interior_field(p, 0.1, 50.0, 0.5, 298e6)
    m = np.sqrt(np.conj(50.0 - 1j * 0.5 / (EO * w)))
    E1, _, _ = mie.bh_fields(p, 0.1, m, w / C0, 40)
    assert np.allclose(Em, np.conj(E1), rtol=1e-12)
    assert np.all(mie.mie_interior_field(np.array([[0.2, 0.0, 0.0]]), 0.1, 50.0, 0.5, 298e6) == 0)
