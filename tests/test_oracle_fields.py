"""The plane-wave excitation pinned by quadrature: PlaneWave_Excitation.m:39-95
averages E0 exp(-j k.r) over each voxel in closed form; a 16-point
Gauss-Legendre rule per axis over the voxel reproduces it (the integrand is
smooth on the voxel), for an oblique wave vector and for one with a zero
component (the factor dx branch). The package's host routine equals the
oracle's."""
import numpy as np

from oracle import fields


def _grid(n, dx):
    a = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(a, a + 0.3 * dx, a - 0.2 * dx, indexing='ij')
    return np.stack([X, Y, Z], axis=-1)


def _quadrature_average(r0, dx, k, Eo):
    u, w = np.polynomial.legendre.leggauss(16)
    pts = u * dx / 2
    tot = 0.0
    for a, wa in zip(pts, w):
        for b, wb in zip(pts, w):
            for c, wc in zip(pts, w):
                p = r0 + np.array([a, b, c])
                tot = tot + wa * wb * wc * np.exp(-1j * np.dot(k, p))
    return Eo * tot / 8.0            # the weights sum to 2 per axis


def test_plane_wave_voxel_average_by_quadrature():
    dx = 0.02
    r = _grid(3, dx)
    em = fields.em_constants(298e6)
    for k in (em['ko'] * np.array([0.3, 0.5, np.sqrt(1 - 0.34)]), np.array([0.0, 0.0, 40.0])):
        Eo = np.array([1.0, 0.5j, -0.25])
        E, H = fields.plane_wave_excitation(r, k, em['omega_mu'], Eo)
        Ho = np.cross(k, Eo) / em['omega_mu']
        for idx in [(0, 0, 0), (1, 2, 0), (2, 1, 2)]:
            avg = _quadrature_average(r[idx], dx, k, 1.0)     # voxel average of exp(-j k.r)
            assert np.allclose(E[idx], Eo * avg, rtol=1e-12, atol=0), (k, idx)
            assert np.allclose(H[idx], Ho * avg, rtol=1e-12, atol=0), (k, idx)


def test_package_plane_wave_matches_oracle():
    from marie_amd import vie
    dx = 0.01
    r = _grid(5, dx)
    em = fields.em_constants(128e6)
    k = em['ko'] * np.array([0.0, 0.6, 0.8])
    E0, H0 = fields.plane_wave_excitation(r, k, em['omega_mu'], np.array([1.0, 0, 0]))
    E1, H1 = vie.PlaneWave_Excitation(r, k, em['omega_mu'], [1, 0, 0])
    assert np.allclose(E1, E0, rtol=1e-14, atol=0) and np.allclose(H1, H0, rtol=1e-14, atol=0)
