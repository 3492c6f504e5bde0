"""Adjoint eps_r / sigma gradients on the device.

- the adjoint GMRES (A^H through mv_AN's 'transp' branch) against
  oracle.krylov.pgmres on the same operator (same residual history);
- the per-voxel gradient through the torch.autograd harness against the
  oracle's adjoint gradient (itself pinned by finite differences on the CPU);
- the harness end to end against central finite differences of device
  forward solves, for an objective of the interior field (coefficient of
  variation of |E_x| over a region plus an L2 misfit), where eps_r and sigma
  also enter explicitly through E = J / tau: relative error <= 1e-5.
"""
import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import fields, krylov, mvp
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def _setup(n=8, dx=6.25e-3, f=298.2e6, seed=11):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    body = X ** 2 + Y ** 2 + 1.2 * Z ** 2 < (0.45 * n * dx) ** 2
    eps = np.where(body, 20 + 40 * rng.random(body.shape), 1.0)
    sig = np.where(body, 0.1 + 1.5 * rng.random(body.shape), 0.0)
    em = fields.em_constants(f)
    E1, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    E2, _ = fields.plane_wave_excitation(r, [em['ko'], 0, 0], em['omega_mu'], [0, 1, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    return dict(n=n, dx=dx, f=f, body=body, eps=eps, sig=sig, E=[E1, E2], fN=fN, rng=rng, r=r)


@pytest.fixture(scope='module')
def case():
    return _setup()


def _problem(case, nrhs, tol=1e-12):
    from marie_amd import adjoint, lib
    n = case['n']
    op = lib.Operator(n, n, n)
    op.set_fN(case['fN'])
    return adjoint.JVIEProblem(op, case['body'], case['E'][:nrhs], case['f'], case['dx'], tol=tol)


def _params(case):
    import torch
    from marie_amd import adjoint
    e = torch.tensor(adjoint.to_voxels(case['eps']), dtype=torch.float64, device='cuda:0', requires_grad=True)
    s = torch.tensor(adjoint.to_voxels(case['sig']), dtype=torch.float64, device='cuda:0', requires_grad=True)
    return e, s


def test_adjoint_gmres_parity(case):
    from marie_amd import lib
    print("case adjoint gmres parity", flush=True)
    em = fields.em_constants(case['f'])
    e_r = case['eps'] - 1j * case['sig'] / (em['eo'] * em['omega'])
    Mr, Mc, Gram, idx, b = fields.jvie_setup(case['E'][0], e_r, case['dx'], case['f'], form=2)
    g = case['rng'].standard_normal(b.size) + 1j * case['rng'].standard_normal(b.size)
    AH = lambda v: mvp.mv_AN(v, case['fN'], Mr, Mc, Gram, 'transp', idx)
    yr, flr, rr, itr, rvr = krylov.pgmres(AH, g, 20, 1e-9, 100)
    n = case['n']
    op = lib.Operator(n, n, n)
    op.set_fN(case['fN'])
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    import torch
    gd = torch.from_numpy(g).to('cuda:0')
    yd = torch.zeros_like(gd)
    fl, rel, it, rv = op.gmres_adjoint_device(gd.data_ptr(), yd.data_ptr(), restart=20, tol=1e-9, maxit=100)
    assert fl == flr == 0 and it == itr
    assert np.max(np.abs(rv - rvr)) <= 1e-9
    y = yd.cpu().numpy()
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) <= 1e-8


def test_gradient_matches_oracle(case):
    import torch
    from marie_amd import adjoint
    print("case adjoint vs oracle", flush=True)
    prob = _problem(case, 1)
    e, s = _params(case)
    X = adjoint.jvie_solve(e, s, prob)
    rng = np.random.default_rng(3)
    w = torch.tensor(rng.random(X.shape[1]), device='cuda:0')
    c = torch.tensor(rng.standard_normal(X.shape[1]) + 1j * rng.standard_normal(X.shape[1]), device='cuda:0')
    c = c * (torch.linalg.norm(X.detach()) / torch.linalg.norm(c))
    J = torch.sum(w * X[0].abs() ** 2) + torch.real(torch.vdot(c, X[0]))
    J.backward()
    # oracle: same objective, g = w x + c / 2
    x, idx, Mc, einc, Gram, em = oadj.forward(case['eps'], case['sig'], case['E'][0], case['dx'], case['f'],
                                              case['fN'], tol=1e-12)
    assert np.linalg.norm(X[0].detach().cpu().numpy() - x) / np.linalg.norm(x) <= 1e-9
    g = w.cpu().numpy() * x + c.cpu().numpy() / 2
    lam = oadj.adjoint_state(g, case['fN'], Mc, Gram, idx, tol=1e-12)
    ge, gs = oadj.material_gradient(lam, x, einc, case['fN'], Mc, Gram, idx, em['omega'] * em['eo'])
    ge, gs = adjoint.to_voxels(ge), adjoint.to_voxels(gs)
    got_e, got_s = e.grad.cpu().numpy(), s.grad.cpu().numpy()
    assert np.max(np.abs(got_e - ge)) <= 1e-8 * np.max(np.abs(ge))
    assert np.max(np.abs(got_s - gs)) <= 1e-8 * np.max(np.abs(gs))
    body = adjoint.to_voxels(case['body'])
    assert np.all(got_e[~body] == 0) and np.all(got_s[~body] == 0)


def test_gradient_with_refined_solves(case):
    """The autograd gradient with complex64-refined forward and adjoint solves
    (GMRES mode 3, block-orthogonalised cycles, FP64 refinement to 1e-11)
    equals the oracle's FP64 adjoint gradient to 1e-8 of its largest entry."""
    import torch
    from marie_amd import adjoint, lib
    print("case adjoint with refined solves", flush=True)
    n = case['n']
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(case['fN'])
    prob = adjoint.JVIEProblem(op, case['body'], case['E'][:1], case['f'], case['dx'], tol=1e-11, restart=40,
                               mode=3)
    e, s = _params(case)
    X = adjoint.jvie_solve(e, s, prob)
    rng = np.random.default_rng(4)
    w = torch.tensor(rng.random(X.shape[1]), device='cuda:0')
    J = torch.sum(w * X[0].abs() ** 2)
    J.backward()
    x, idx, Mc, einc, Gram, em = oadj.forward(case['eps'], case['sig'], case['E'][0], case['dx'], case['f'],
                                              case['fN'], tol=1e-12)
    assert np.linalg.norm(X[0].detach().cpu().numpy() - x) / np.linalg.norm(x) <= 1e-8
    lam = oadj.adjoint_state(w.cpu().numpy() * x, case['fN'], Mc, Gram, idx, tol=1e-12)
    ge, gs = oadj.material_gradient(lam, x, einc, case['fN'], Mc, Gram, idx, em['omega'] * em['eo'])
    ge, gs = adjoint.to_voxels(ge), adjoint.to_voxels(gs)
    assert np.max(np.abs(e.grad.cpu().numpy() - ge)) <= 1e-8 * np.max(np.abs(ge))
    assert np.max(np.abs(s.grad.cpu().numpy() - gs)) <= 1e-8 * np.max(np.abs(gs))


def _field_objective(prob, e, s, target, roi):
    """CV of |E_x| over the region + misfit to a target field (all excitations)."""
    import torch
    from marie_amd import adjoint
    X = adjoint.jvie_solve(e, s, prob)
    E = adjoint.interior_field(X, e, s, prob)          # nrhs x nidx
    nb = E.shape[1] // 3
    ex = E[:, :nb].abs()[:, roi]
    cv = torch.std(ex, dim=1) / torch.mean(ex, dim=1)
    misfit = torch.sum((E - target).abs() ** 2) / torch.sum(target.abs() ** 2)
    return torch.sum(cv) + 0.5 * misfit


def test_autograd_matches_finite_differences(case):
    import torch
    print("case adjoint finite differences", flush=True)
    prob = _problem(case, 2)
    e, s = _params(case)
    nb = prob.idx.size // 3
    rng = np.random.default_rng(7)
    roi = torch.tensor(rng.random(nb) < 0.5, device='cuda:0')
    with torch.no_grad():
        from marie_amd import adjoint
        X0 = adjoint.jvie_solve(e, s, prob)
        E0 = adjoint.interior_field(X0, e, s, prob)
        target = E0 * (1 + 0.3 * torch.tensor(rng.standard_normal(E0.shape), device='cuda:0'))
    J = _field_objective(prob, e, s, target, roi)
    J.backward()
    Jv = float(J.detach())
    ge, gs = e.grad.cpu().numpy(), s.grad.cpu().numpy()
    body = np.nonzero(adjoint.to_voxels(case['body']))[0]
    for v in rng.choice(body, 6, replace=False):
        for which, grad in (('eps', ge), ('sig', gs)):
            base = (e if which == 'eps' else s).detach()
            h = 1e-4 * float(base[v])
            vals = []
            for sgn in (1, -1):
                p = base.clone()
                p[v] += sgn * h
                with torch.no_grad():
                    ee, ss = (p, s.detach()) if which == 'eps' else (e.detach(), p)
                    vals.append(float(_field_objective(prob, ee, ss, target, roi)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grad[v]) <= 1e-5 * abs(fd) + 1e-9 * abs(Jv), (which, v, fd, grad[v])


def test_b1plus_objective_gradient():
    """Coil-driven objective of the differentiation workload in miniature: four
    loop ports driven in circular polarisation, B1+ = (Bx + j By)/2 of the
    total field B = mu0 (Hinc + K J / Gram); J = coefficient of variation of
    |B1+| over a central region + L2 misfit to a seeded target. Autograd
    through the adjoint solve and the K product against central differences."""
    import torch
    from marie_amd import adjoint, coil, geometry, lib, vie
    print("case B1+ objective", flush=True)
    n, dx, f = 8, 8e-3, 128e6
    r = geometry.cell_grid(n, n, n, dx)
    em = vie.EMconstants(f)
    R = np.linalg.norm(r, axis=-1)
    body = R < 0.03
    rng = np.random.default_rng(9)
    eps = np.where(body, 30 + 30 * rng.random(body.shape), 1.0)
    sig = np.where(body, 0.2 + rng.random(body.shape), 0.0)
    ports = coil.cylinder_loop_array(4, 0.08, 0.06, 0.03, seg_per_side=4)
    Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    op = lib.Operator(n, n, n)
    op.assemble_N(dx, em['ko'])
    op.assemble_K(dx, em['ko'])
    prob = adjoint.JVIEProblem(op, body, Es, f, dx, tol=1e-12, Hinc=Hs)
    w = torch.tensor(np.exp(-2j * np.pi * np.arange(4) / 4), device='cuda:0')
    nb = prob.idx.size // 3
    roi = torch.tensor(np.linalg.norm(r.reshape(-1, 3, order='F')[prob.idx[:nb]], axis=1) < 0.02, device='cuda:0')

    def objective(e, s, target):
        X = adjoint.jvie_solve(e, s, prob)
        b1 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X, prob), prob), dim=0)
        a = b1.abs()[roi]
        cv = torch.std(a) / torch.mean(a)
        return cv + torch.sum((b1 - target).abs() ** 2) / torch.sum(target.abs() ** 2)

    e = torch.tensor(adjoint.to_voxels(eps), device='cuda:0', requires_grad=True)
    s = torch.tensor(adjoint.to_voxels(sig), device='cuda:0', requires_grad=True)
    with torch.no_grad():
        X0 = adjoint.jvie_solve(e, s, prob)
        b0 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X0, prob), prob), dim=0)
        target = b0 * (1 + 0.2 * torch.tensor(rng.standard_normal(b0.shape), device='cuda:0'))
    J = objective(e, s, target)
    J.backward()
    Jv = float(J.detach())
    ge, gs = e.grad.cpu().numpy(), s.grad.cpu().numpy()
    vox = np.nonzero(adjoint.to_voxels(body))[0]
    for v in rng.choice(vox, 5, replace=False):
        for which, grad in (('eps', ge), ('sig', gs)):
            base = (e if which == 'eps' else s).detach()
            h = 1e-4 * float(base[v])
            vals = []
            for sgn in (1, -1):
                p = base.clone()
                p[v] += sgn * h
                with torch.no_grad():
                    ee, ss = (p, s.detach()) if which == 'eps' else (e.detach(), p)
                    vals.append(float(objective(ee, ss, target)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grad[v]) <= 1e-5 * abs(fd) + 1e-9 * abs(Jv), (which, v, fd, grad[v])


def test_b1plus_gradient_64_voxels():
    """The differentiation workload at reduced size: a seeded head phantom on a
    24^3 grid over the 0.192 m domain, the 8 ports of a two-ring loop coil in
    circular polarisation, objective = CV of |B1+| over a central ellipsoid +
    L2 misfit to a seeded target. Autograd (adjoint solves + K product +
    per-voxel reduction) against central differences on 64 random voxels
    (32 eps_r, 32 sigma)."""
    import torch
    from marie_amd import adjoint, coil, geometry, lib, vie
    print("case B1+ gradient, 64 voxels", flush=True)
    n, f = 24, 298e6
    dx = 0.192 / n
    r, eps, sig, _ = geometry.head_phantom(n, n, n, dx, seed=5, scale=0.85)
    em = vie.EMconstants(f)
    body = eps > 1.0
    ports = coil.cylinder_loop_array(8, 0.15, 0.1, 0.05, seg_per_side=4, rings=2, length=0.2)
    Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    op = lib.Operator(n, n, n)
    op.assemble_N(dx, em['ko'])
    op.assemble_K(dx, em['ko'])
    prob = adjoint.JVIEProblem(op, body, Es, f, dx, tol=1e-11, Hinc=Hs)
    nb = prob.idx.size // 3
    w = torch.tensor(np.exp(-2j * np.pi * np.arange(8) / 4), device='cuda:0')   # CP within each ring
    pts = r.reshape(-1, 3, order='F')[prob.idx[:nb]]
    roi = torch.tensor(np.sum((pts / np.array([0.04, 0.035, 0.05])) ** 2, axis=1) < 1.0, device='cuda:0')
    assert int(roi.sum()) > 20

    def objective(e, s, target):
        X = adjoint.jvie_solve(e, s, prob)
        b1 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X, prob), prob), dim=0)
        a = b1.abs()[roi]
        return torch.std(a) / torch.mean(a) + torch.sum((b1 - target).abs() ** 2) / torch.sum(target.abs() ** 2)

    rng = np.random.default_rng(12)
    e = torch.tensor(adjoint.to_voxels(eps), device='cuda:0', requires_grad=True)
    s = torch.tensor(adjoint.to_voxels(sig), device='cuda:0', requires_grad=True)
    with torch.no_grad():
        X0 = adjoint.jvie_solve(e, s, prob)
        b0 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X0, prob), prob), dim=0)
        target = b0 * (1 + 0.2 * torch.tensor(rng.standard_normal(b0.shape), device='cuda:0'))
    J = objective(e, s, target)
    J.backward()
    Jv = float(J.detach())
    ge, gs = e.grad.cpu().numpy(), s.grad.cpu().numpy()
    vox = np.nonzero(adjoint.to_voxels(body))[0]
    picks = rng.choice(vox, 64, replace=False)
    worst = 0.0
    resolved = 0
    for t, v in enumerate(picks):
        which, grad = ('eps', ge) if t < 32 else ('sig', gs)
        base = (e if which == 'eps' else s).detach()
        # central differences with steps h and h/2 (h = 1e-3 of the value),
        # Richardson-combined so the truncation is O(h^4); the solves' 1e-11
        # tolerance leaves J accurate to ~1e-11 relative, a floor of about
        # 6e-11 |J| / h on the combined quotient
        h = 1e-3 * float(base[v])

        def quotient(step):
            vals = []
            for sgn in (1, -1):
                p = base.clone()
                p[v] += sgn * step
                with torch.no_grad():
                    ee, ss = (p, s.detach()) if which == 'eps' else (e.detach(), p)
                    vals.append(float(objective(ee, ss, target)))
            return (vals[0] - vals[1]) / (2 * step)

        fd = (4 * quotient(h / 2) - quotient(h)) / 3
        err = abs(fd - grad[v])
        worst = max(worst, err / max(abs(fd), 1e-300))
        floor = 6e-11 * abs(Jv) / h
        assert err <= 1e-5 * abs(fd) + floor, (which, v, fd, grad[v])
        if abs(fd) > 1e3 * floor:
            resolved += 1
    print("  worst relative deviation over 64 voxels %.2e; %d gradients resolved to 1e-5 above the "
          "difference quotient's noise" % (worst, resolved), flush=True)
    assert resolved >= 16
