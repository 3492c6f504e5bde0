"""Voxel domains, homogeneous shapes and the synthetic head-like phantom.

Host-side input synthesis (the reference's src_geometry is MATLAB host code
too): generatedomain.m:40-94 / grid3d.m for the grid, homogenellipse.m:79-88
for the voxel-centre inside test.
"""
import numpy as np


def generatedomain(res, x, y=None, z=None):
    """r (L x M x N x 3) of voxel centres (generatedomain.m:40-94).

    A scalar extent e means [-e, e]; the grid has 2*ceil(D/2/res)+1 points.
    """
    def _axis(v):
        v = np.atleast_1d(np.asarray(v, float))
        if v.size == 1:
            v = np.array([-v[0], v[0]])
        D = v.max() - v.min()
        n = int(np.ceil(D / 2 / res))
        C = (v.max() + v.min()) / 2
        return C + res * np.arange(-n, n + 1)
    xs = _axis(x)
    ys = _axis(x if y is None else y)
    zs = _axis(x if z is None else z)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    return np.stack([X, Y, Z], axis=-1)


def cell_grid(L, M, N, dx, center=(0.0, 0.0, 0.0)):
    """Cell-centred grid x_i = (i - (L-1)/2) dx + c."""
    xs = (np.arange(L) - (L - 1) / 2) * dx + center[0]
    ys = (np.arange(M) - (M - 1) / 2) * dx + center[1]
    zs = (np.arange(N) - (N - 1) / 2) * dx + center[2]
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    return np.stack([X, Y, Z], axis=-1)


def inside_ellipsoid(r, center, semi):
    """Voxel-centre test of homogenellipse.m:79-80 (strict inequality)."""
    c = np.asarray(center, float)
    s = np.asarray(semi, float)
    q = sum(((r[..., a] - c[a]) / s[a]) ** 2 for a in range(3))
    return q < 1.0


def sphere_body(r, center, radius, eps_r, sigma_e, rho=1000.0):
    """homogensphere.m: epsilon_r, sigma_e, rho maps and the inside index."""
    inside = inside_ellipsoid(r, center, (radius, radius, radius))
    epsilon_r = np.where(inside, eps_r, 1.0)
    sigma = np.where(inside, sigma_e, 0.0)
    dens = np.where(inside, rho, 0.0)
    return inside, epsilon_r, sigma, dens


PHANTOM_SEMI = (0.09, 0.08, 0.11)   # outer ellipsoid semi-axes (m)
PHANTOM_SHELLS = (1.0, 0.85, 0.7)   # shell scale factors, outer to inner


def head_phantom(L, M, N, dx, seed=0, scale=1.0):
    """Three nested ellipsoids with per-shell eps_r ~ U[20,80], sigma ~ U[0.1,2] S/m.

    Returns (r, epsilon_r, sigma_e, params). `scale` shrinks the whole phantom
    (used when the grid is smaller than the 0.22 m outer z-diameter).
    """
    rng = np.random.default_rng(seed)
    eps = rng.uniform(20.0, 80.0, size=3)
    sig = rng.uniform(0.1, 2.0, size=3)
    r = cell_grid(L, M, N, dx)
    epsilon_r = np.ones((L, M, N))
    sigma_e = np.zeros((L, M, N))
    for s, e, g in zip(PHANTOM_SHELLS, eps, sig):
        semi = np.array(PHANTOM_SEMI) * s * scale
        ins = inside_ellipsoid(r, (0, 0, 0), semi)
        epsilon_r[ins] = e
        sigma_e[ins] = g
    return r, epsilon_r, sigma_e, dict(seed=seed, eps=eps.tolist(), sigma=sig.tolist(), scale=scale)
