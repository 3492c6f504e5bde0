"""Oracle: weakly singular face-pair integrals of exp(-jkR)/R over unit-cell squares.

The reference evaluates five 4D integrals (coincident ST, edge-adjacent
coplanar EAc / orthogonal EAo, vertex-adjacent coplanar VAc / orthogonal VAo)
with DIRECTFN or DEMCEM (singular_st.m:12-20, singular_ea.m:61-76,
singular_va.m:131-147), each as a sum of four triangle-pair integrals. The
value is  I = int_{S_p} int_{S_q} exp(-j k |r - r'|) / |r - r'| dS dS'  over two
squares of side a, in absolute units (m^3), the same quantity the regular
cubature fills in (cubature_nop_sym_near.m:88).

This module's own method (test infrastructure; the HIP library implements the
same derivation on the device):

* Both squares are axis-aligned, so the 4D integral is a convolution and
  reduces to an integral over the difference vector t = r' - r with a
  piecewise-linear density (the overlap length along each axis).
* Coplanar cases (ST, EAc, VAc) become 2D integrals over t; the sub-square
  that touches t = 0 is integrated in polar coordinates about t = 0, whose
  Jacobian r cancels the 1/r singularity; the other sub-squares are regular.
* Orthogonal cases (EAo, VAo) become 3D integrals over (u, v, s); the sub-cube
  touching the origin is split into three pyramids with apex at the origin,
  whose Jacobian (a lambda)^2 cancels the singularity; the rest is regular.

Every resulting integrand is analytic on its domain, so tensor Gauss-Legendre
converges geometrically; 24 points per direction reach double precision.
"""
import numpy as np

NQ = 24


def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0


def _polar_square(f, a, n):
    """int over [0,a]^2 of F(r, th) dr dth, with the r-Jacobian already in F.

    f(r, c, s) is the integrand times r (so that the singular 1/r is gone).
    Splits the square into the triangles th in [0,pi/4] and [pi/4,pi/2].
    """
    x, w = _gl01(n)
    tot = 0.0 + 0.0j
    for lo, hi, rmax in ((0.0, np.pi / 4, lambda c, s: a / c),
                         (np.pi / 4, np.pi / 2, lambda c, s: a / s)):
        th = lo + (hi - lo) * x
        wt = (hi - lo) * w
        c, s = np.cos(th), np.sin(th)
        R = rmax(c, s)                                  # n
        r = R[:, None] * x[None, :]                     # n x n
        wr = R[:, None] * w[None, :]
        tot += np.sum(wt[:, None] * wr * f(r, c[:, None], s[:, None]))
    return tot


def _box2(f, x0, x1, y0, y1, n):
    x, w = _gl01(n)
    t1 = x0 + (x1 - x0) * x
    t2 = y0 + (y1 - y0) * x
    W = np.outer(w * (x1 - x0), w * (y1 - y0))
    T1, T2 = np.meshgrid(t1, t2, indexing='ij')
    return np.sum(W * f(T1, T2))


def _box3(f, lims, n):
    x, w = _gl01(n)
    pts, wts = [], []
    for lo, hi in lims:
        pts.append(lo + (hi - lo) * x)
        wts.append(w * (hi - lo))
    A, B, C = np.meshgrid(*pts, indexing='ij')
    W = wts[0][:, None, None] * wts[1][None, :, None] * wts[2][None, None, :]
    return np.sum(W * f(A, B, C))


def _pyr(k, a, n):
    """(lambda, p, q) nodes/weights on [0,1]^3 plus D = sqrt(1+p^2+q^2)."""
    x, w = _gl01(n)
    lam, p, q = np.meshgrid(x, x, x, indexing='ij')
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    D = np.sqrt(1.0 + p * p + q * q)
    E = np.exp(-1j * k * a * lam * D) / D            # exp(-jk rho) * a*lam / rho
    return lam, p, q, W, E


def g(k, r):
    return np.exp(-1j * k * r) / r


def st(k, a, n=NQ):
    """Coincident squares: 4 int_[0,a]^2 g(|t|)(a-t1)(a-t2) dt."""
    f = lambda r, c, s: np.exp(-1j * k * r) * (a - r * c) * (a - r * s)
    return 4.0 * _polar_square(f, a, n)


def eac(k, a, n=NQ):
    """Edge-adjacent coplanar squares [0,a]^2 and [a,2a]x[0,a]."""
    fa = lambda r, c, s: np.exp(-1j * k * r) * (r * c) * (a - r * s)
    A = _polar_square(fa, a, n)
    fb = lambda t1, t2: g(k, np.hypot(t1, t2)) * (2 * a - t1) * (a - t2)
    B = _box2(fb, a, 2 * a, 0.0, a, n)
    return 2.0 * (A + B)


def vac(k, a, n=NQ):
    """Vertex-adjacent coplanar squares [0,a]^2 and [a,2a]^2."""
    fa = lambda r, c, s: np.exp(-1j * k * r) * (r * c) * (r * s)
    A = _polar_square(fa, a, n)
    fb = lambda t1, t2: g(k, np.hypot(t1, t2)) * (2 * a - t1) * t2
    B = _box2(fb, a, 2 * a, 0.0, a, n)
    fc = lambda t1, t2: g(k, np.hypot(t1, t2)) * (2 * a - t1) * (2 * a - t2)
    C = _box2(fc, a, 2 * a, a, 2 * a, n)
    return A + 2.0 * B + C


def eao(k, a, n=NQ):
    """Edge-adjacent orthogonal squares z=0:[0,a]^2 and x=a:[0,a]^2.

    I = 2 int_[0,a]^3 (a - s) g(sqrt(u^2+v^2+s^2)); pyramids with bases
    u=a and v=a carry s = a lam q, the one with base s=a carries s = a lam.
    """
    lam, p, q, W, E = _pyr(k, a, n)
    base = a ** 2 * lam * E                             # g * Jacobian
    Puv = np.sum(W * base * a * (1.0 - lam * q))
    Ps = np.sum(W * base * a * (1.0 - lam))
    return 2.0 * (2.0 * Puv + Ps)


def vao(k, a, n=NQ):
    """Vertex-adjacent orthogonal squares z=0:[0,a]^2 and x=a:[a,2a]x[0,a]."""
    lam, p, q, W, E = _pyr(k, a, n)
    base = a ** 2 * lam * E
    A = 2.0 * np.sum(W * base * a * lam * q) + np.sum(W * base * a * lam)
    fb = lambda u, v, s: g(k, np.sqrt(u * u + v * v + s * s)) * (2 * a - s)
    B = _box3(fb, [(0.0, a), (0.0, a), (a, 2 * a)], n)
    return A + B


def all_terms(k, a, n=NQ):
    """(I_ST, I_EAc, I_EAo, I_VAc, I_VAo) in the order assembly_nop.m:96-98 uses."""
    return st(k, a, n), eac(k, a, n), eao(k, a, n), vac(k, a, n), vao(k, a, n)


def st_static_closed_form(a):
    """k -> 0 limit of the coincident-square integral, a^3[(4/3)(1-sqrt2)+4 ln(1+sqrt2)]."""
    return a ** 3 * ((4.0 / 3.0) * (1.0 - np.sqrt(2.0)) + 4.0 * np.log(1.0 + np.sqrt(2.0)))
