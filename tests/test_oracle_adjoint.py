"""The oracle's adjoint gradient against central finite differences of its
own forward solve (the pin of oracle.adjoint)."""
import numpy as np

from oracle import adjoint, fields
from oracle import operators as oop


def _setup(n=7, dx=6.25e-3, f=298.2e6, seed=5):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    body = X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2
    eps = np.where(body, 20 + 40 * rng.random(body.shape), 1.0)
    sig = np.where(body, 0.1 + 1.5 * rng.random(body.shape), 0.0)
    em = fields.em_constants(f)
    Einc, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    return eps, sig, Einc, dx, f, fN, body, rng


def _objective(x, w, c):
    return float(np.sum(w * np.abs(x) ** 2) + np.real(np.vdot(c, x)))


def test_adjoint_gradient_matches_finite_differences():
    eps, sig, Einc, dx, f, fN, body, rng = _setup()
    x, idx, Mc, einc, Gram, em = adjoint.forward(eps, sig, Einc, dx, f, fN)
    w = rng.random(x.size)
    c = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
    c *= np.linalg.norm(x) / np.linalg.norm(c)
    g = w * x + c / 2                              # dJ/dconj(x)
    lam = adjoint.adjoint_state(g, fN, Mc, Gram, idx)
    omega_eo = em['omega'] * em['eo']
    ge, gs = adjoint.material_gradient(lam, x, einc, fN, Mc, Gram, idx, omega_eo)
    assert np.all(ge[~body] == 0) and np.all(gs[~body] == 0)
    vox = np.argwhere(body)
    pick = vox[rng.choice(len(vox), 5, replace=False)]
    for v in map(tuple, pick):
        for which, grad, base, h in (('eps', ge, eps, 1e-4 * eps[v]), ('sig', gs, sig, 1e-4 * sig[v])):
            vals = []
            for sgn in (1, -1):
                p = base.copy()
                p[v] += sgn * h
                xe, *_ = adjoint.forward(p if which == 'eps' else eps, p if which == 'sig' else sig, Einc, dx, f, fN)
                vals.append(_objective(xe, w, c))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grad[v]) <= 1e-5 * abs(fd) + 1e-12 * abs(_objective(x, w, c)), (which, v, fd, grad[v])


def test_gradient_of_zero_objective_direction_is_zero():
    """g = 0 -> lambda = 0 -> zero gradient; and the gradient is linear in g."""
    eps, sig, Einc, dx, f, fN, body, rng = _setup(n=6)
    x, idx, Mc, einc, Gram, em = adjoint.forward(eps, sig, Einc, dx, f, fN)
    omega_eo = em['omega'] * em['eo']
    g1 = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
    g2 = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
    l1 = adjoint.adjoint_state(g1, fN, Mc, Gram, idx)
    l2 = adjoint.adjoint_state(g2, fN, Mc, Gram, idx)
    l12 = adjoint.adjoint_state(g1 + 3 * g2, fN, Mc, Gram, idx)
    a = adjoint.material_gradient(l1, x, einc, fN, Mc, Gram, idx, omega_eo)
    b = adjoint.material_gradient(l2, x, einc, fN, Mc, Gram, idx, omega_eo)
    ab = adjoint.material_gradient(l12, x, einc, fN, Mc, Gram, idx, omega_eo)
    for k in range(2):
        ref = a[k] + 3 * b[k]
        assert np.max(np.abs(ab[k] - ref)) <= 1e-9 * np.max(np.abs(ref))
