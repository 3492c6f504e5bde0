"""Oracle: interior field of a homogeneous sphere under a plane wave (Mie series).

Used for the config-1 physics check the reference performs in
scripts/vie_solver_mie/RUN_Sphere_VIE_MIE.m:211 (MIE_SERIES.m:19-44,
polar_transmitted_field.m:13-60). Restated from Bohren & Huffman, "Absorption
and Scattering of Light by Small Particles" (1983), eqs. 4.50 and 4.52-4.53,
in their exp(-i w t) convention with incident x-polarised exp(ikz):

  E_1 = sum_n E_n (c_n M_o1n - i d_n N_e1n),  E_n = i^n (2n+1)/(n(n+1)),
  c_n = (i/x) / (j_n(mx) [x h_n(x)]' - h_n(x) [mx j_n(mx)]'),
  d_n = m (i/x) / (m^2 j_n(mx) [x h_n(x)]' - h_n(x) [mx j_n(mx)]')

(mu = mu_1 = 1; the numerators reduce by the Wronskian j_n y_n' - j_n' y_n = 1/x^2).
MARIE uses exp(+j w t), e_r = eps_r - j sigma/(w eps0) and an incident
exp(-jkz): its field is the complex conjugate of the B&H field computed with
m = sqrt(conj(e_r)). Test infrastructure only.
"""
import numpy as np
from scipy import special


def _pi_tau(nmax, mu):
    pi = np.zeros((nmax + 1,) + mu.shape)
    tau = np.zeros_like(pi)
    pi[1] = 1.0
    for n in range(2, nmax + 1):
        pi[n] = (2 * n - 1) / (n - 1) * mu * pi[n - 1] - n / (n - 1) * pi[n - 2]
    for n in range(1, nmax + 1):
        tau[n] = n * mu * pi[n] - (n + 1) * pi[n - 1]
    return pi, tau


def mie_coefficients(x, m, nmax):
    n = np.arange(1, nmax + 1)
    jx = special.spherical_jn(n, x)
    yx = special.spherical_yn(n, x)
    hx = jx + 1j * yx
    djx = special.spherical_jn(n, x, derivative=True)
    dyx = special.spherical_yn(n, x, derivative=True)
    dxh = hx + x * (djx + 1j * dyx)                  # [x h_n(x)]'
    mx = m * x
    jmx = special.spherical_jn(n, mx)
    dmxj = jmx + mx * special.spherical_jn(n, mx, derivative=True)   # [mx j_n(mx)]'
    c = (1j / x) / (jmx * dxh - hx * dmxj)
    d = m * (1j / x) / (m * m * jmx * dxh - hx * dmxj)
    return c, d


def _fields_bh(r, k, nmax, coef_M, coef_N, radial, mk):
    """sum_n E_n (coef_M[n] M_o1n - i coef_N[n] N_e1n) with radial function `radial`
    (spherical_jn or the first-kind Hankel function) of argument mk*r, B&H (4.50)."""
    pts = np.asarray(r, float)
    X, Y, Z = pts[..., 0], pts[..., 1], pts[..., 2]
    R = np.sqrt(X * X + Y * Y + Z * Z)
    Rs = np.maximum(R, 1e-12)
    th = np.arccos(np.clip(Z / Rs, -1, 1))
    ph = np.arctan2(Y, X)
    mu = np.cos(th)
    st = np.sin(th)
    pi, tau = _pi_tau(nmax, mu)
    rho = mk * Rs
    Er = np.zeros(R.shape, complex)
    Et = np.zeros(R.shape, complex)
    Ep = np.zeros(R.shape, complex)
    cp, sp = np.cos(ph), np.sin(ph)
    for n in range(1, nmax + 1):
        En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
        zn, dzn = radial(n, rho)
        drz = zn + rho * dzn                                   # [rho z_n(rho)]'
        Mt = cp * pi[n] * zn
        Mp = -sp * tau[n] * zn
        Nr = cp * n * (n + 1) * st * pi[n] * zn / rho
        Nt = cp * tau[n] * drz / rho
        Np_ = -sp * pi[n] * drz / rho
        Er += En * (-1j * coef_N[n - 1] * Nr)
        Et += En * (coef_M[n - 1] * Mt - 1j * coef_N[n - 1] * Nt)
        Ep += En * (coef_M[n - 1] * Mp - 1j * coef_N[n - 1] * Np_)
    Ex = st * cp * Er + mu * cp * Et - sp * Ep
    Ey = st * sp * Er + mu * sp * Et + cp * Ep
    Ez = mu * Er - st * Et
    return np.stack([Ex, Ey, Ez], axis=-1), R


def _jn(n, z):
    return special.spherical_jn(n, z), special.spherical_jn(n, z, derivative=True)


def _h1n(n, z):
    return (special.spherical_jn(n, z) + 1j * special.spherical_yn(n, z),
            special.spherical_jn(n, z, derivative=True) + 1j * special.spherical_yn(n, z, derivative=True))


def scattering_coefficients(x, m, nmax):
    """a_n, b_n of B&H (4.53) with mu = mu_1 = 1."""
    n = np.arange(1, nmax + 1)
    jx, djx = _jn(n, x)
    hx, dhx = _h1n(n, x)
    mx = m * x
    jmx, djmx = _jn(n, mx)
    dxj = jx + x * djx
    dxh = hx + x * dhx
    dmxj = jmx + mx * djmx
    a = (m * m * jmx * dxj - jx * dmxj) / (m * m * jmx * dxh - hx * dmxj)
    b = (jmx * dxj - jx * dmxj) / (jmx * dxh - hx * dmxj)
    return a, b


def bh_fields(r, a, m, k, nmax):
    """(E_inside, E_scattered, E_incident) in B&H's exp(-i w t) convention at points r."""
    c, d = mie_coefficients(k * a, m, nmax)
    an, bn = scattering_coefficients(k * a, m, nmax)
    E1, R = _fields_bh(r, k, nmax, c, d, _jn, m * k)
    # scattered: E_s = sum E_n (i a_n N_e1n^(3) - b_n M_o1n^(3))  (B&H 4.45)
    Es, _ = _fields_bh(r, k, nmax, -bn, -an, _h1n, k)
    pts = np.asarray(r, float)
    Ei = np.zeros(pts.shape, complex)
    Ei[..., 0] = np.exp(1j * k * pts[..., 2])
    return E1, Es, Ei


def mie_interior_field(r, a, eps_r, sigma, freq, nmax=None):
    """E inside a sphere of radius a centred at 0 (zero outside), MARIE's convention.

    r: (..., 3) points; unit x-polarised incident field travelling along +z.
    """
    mu0 = 4 * np.pi * 1e-7
    c0 = 299792458.0
    eo = 1.0 / c0 ** 2 / mu0
    w = 2 * np.pi * freq
    k = w / c0
    e_r = eps_r - 1j * sigma / (eo * w)
    m = np.sqrt(np.conj(e_r))
    x = k * a
    if nmax is None:
        nmax = int(np.ceil(abs(m) * x + 4 * (abs(m) * x) ** (1 / 3) + 10))
    c, d = mie_coefficients(x, m, nmax)
    E, R = _fields_bh(r, k, nmax, c, d, _jn, m * k)
    E[R > a] = 0.0
    return np.conj(E)
