"""The precision, batching and differentiation workloads at their full sizes.

* complex64-refined solves: the 128^3 head phantom at 1.5 mm, 298 MHz, GMRES
  with complex64 products and basis inside FP64 refinement (mode 2, classical
  Gram-Schmidt cycles of 50; mode 3, block-orthogonalised cycles of 128, the
  benchmark's refined solve). Each true residual is recomputed with the
  oracle's FP64 mv_AN on the CPU (pgmres.m:175 defines the flag by that
  residual), and each error against a 1e-10 FP64 solve is held to the FP64
  1e-6 solve's own error level.
* batch: seeded phantoms of 64^3, 80^3, 96^3 and 128^3 voxels over a 0.192 m
  domain at 128 and 298 MHz, each driven by the 16 ports of a cylinder loop
  coil, solved together by marie_vie_block_gmres_batch, host-driven and with
  device-driven block steps (each step a HIP graph, one host thread for all
  problems); every column's true residual is recomputed, and two ports per
  problem are compared with single GMRES solves (the per-port loops of
  VSIE_Solver.m:387-407 the batch replaces).
* differentiation: the 128^3 phantom, the 8 ports of a two-ring coil in
  circular polarisation, objective = coefficient of variation of |B1+| over a
  central ellipsoid + L2 misfit to a seeded target; the autograd gradient
  (adjoint solves + K product + per-voxel reduction) against Richardson
  central differences re-solved to 2e-12 on random body voxels.
"""
import numpy as np
import pytest

from oracle import mvp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _phantom(L, M, N, dx, f, scale, seed=0):
    from marie_amd import geometry, vie
    r, eps, sig, _ = geometry.head_phantom(L, M, N, dx, seed=seed, scale=scale)
    em = vie.EMconstants(f)
    e_r = eps - 1j * sig / (em['eo'] * em['omega'])
    return r, eps, sig, e_r, em


def test_config3_complex64_refined_solve():
    from marie_amd import lib, vie
    n, dx, f = 128, 1.5e-3, 298e6
    print("case complex64-refined solve, 128^3 phantom, 298 MHz", flush=True)
    r, _, _, e_r, em = _phantom(n, n, n, dx, f, 0.85)
    Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    Mr, Mc, Gram, idx1, b = vie.jvie_system(Einc, e_r, r, f, form=2)
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.assemble_N(dx, em['ko'])
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx1, one_based=True)
    xref, fl, rel_ref, _, _ = op.gmres(b, restart=50, tol=1e-10, maxit=400)
    assert fl == 0
    x64, fl64, rel64, _, _ = op.gmres(b, restart=50, tol=1e-6, maxit=200)
    assert fl64 == 0 and rel64 <= 1e-6
    fN = op.get_fN()
    e64 = _rel(x64, xref)
    for mode, restart in ((2, 50), (3, 128)):
        xm, flm, relm, itm, rvm = op.gmres(b, restart=restart, tol=1e-6, maxit=200, mode=mode, inner_tol=1e-7)
        assert flm == 0 and relm <= 1e-6
        # the true residual of the refined solution by the oracle's FP64 product
        res = b - mvp.mv_AN(xm, fN, Mr, Mc, Gram, 'notransp', idx1 - 1, workers=16)
        true_rel = np.linalg.norm(res) / np.linalg.norm(b)
        em2 = _rel(xm, xref)
        print("  refined mode %d: %d cycles, relres %.2e (oracle %.2e); error vs 1e-10 FP64 %.2e, FP64 1e-6 "
              "solve %.2e" % (mode, itm[1], relm, true_rel, em2, e64), flush=True)
        assert true_rel <= 1e-6
        assert em2 <= 1.2 * e64


def test_config4_batch_bodies_frequencies_ports():
    import torch
    from marie_amd import coil, lib, vie
    print("case batch: 64/80/96/128^3 phantoms x 128/298 MHz x 16 ports", flush=True)
    dev = torch.device('cuda:0')
    ports = coil.cylinder_loop_array(16, 0.15, 0.1, 0.05, seg_per_side=8, rings=2, length=0.2)
    probs, streams = [], []
    for nb, seed in ((64, 1), (80, 2), (96, 3), (128, 4)):
        dxb = 0.192 / nb
        for fb in (128e6, 298e6):
            rb, _, _, e_rb, emb = _phantom(nb, nb, nb, dxb, fb, 0.85, seed=seed)
            Es, _ = coil.CCL_Excitation(rb, emb['ko'], 1.0, ports)
            Mrb, Mcb, Gb, idxb, v0 = vie.jvie_system(Es[0], e_rb, rb, fb, form=2)
            fac = np.tile((Gb * 1j * emb['omega'] * emb['eo'] * (e_rb - 1) / e_rb).ravel(order='F'), 3)[idxb - 1]
            cols = np.stack([fac * Ep.ravel(order='F')[idxb - 1] for Ep in Es])
            assert _rel(cols[0], v0) <= 1e-12
            opb = lib.Operator(nb, nb, nb)
            st = torch.cuda.Stream(device=dev)          # each problem on its own stream
            opb.set_stream(st.cuda_stream)
            streams.append(st)
            opb.assemble_N(dxb, emb['ko'])
            opb.set_material(Mrb, Mcb, Gb)
            opb.set_index(idxb, one_based=True)
            probs.append((nb, fb, opb, torch.from_numpy(np.ascontiguousarray(cols)).to(dev)))
    Xs = [torch.zeros_like(p[3]) for p in probs]
    torch.cuda.synchronize()
    fls, rrs, its = lib.block_gmres_batch([p[2] for p in probs], [p[3].data_ptr() for p in probs],
                                          [x.data_ptr() for x in Xs], [16] * len(probs), restart=20, tol=1e-6,
                                          maxit=100, mode=0, inner_tol=1e-7, use_graph=True)
    torch.cuda.synchronize()
    # the same batch with device-driven block steps (each step a HIP graph),
    # all problems driven from one host thread
    Xd = [torch.zeros_like(p[3]) for p in probs]
    torch.cuda.synchronize()
    fld, rrd, itd = lib.block_gmres_batch([p[2] for p in probs], [p[3].data_ptr() for p in probs],
                                          [x.data_ptr() for x in Xd], [16] * len(probs), restart=20, tol=1e-6,
                                          maxit=100, mode=0, inner_tol=1e-7, use_graph=True, device_steps=True)
    torch.cuda.synchronize()
    for (nb, fb, opb, Bb), X, fl, rr, Xdv, fd, rd in zip(probs, Xs, fls, rrs, Xd, fld, rrd):
        assert int(fd.max()) == 0 and float(rd.max()) <= 1e-6, (nb, fb)
        Bh, Xh = Bb.cpu().numpy(), Xdv.cpu().numpy()
        AX = opb.apply(Xh.T)
        worst = max(_rel(AX[:, q], Bh[q]) for q in range(16))
        assert worst <= 1e-6, (nb, fb, worst)
        assert _rel(Xh, X.cpu().numpy()) <= 1e-5, (nb, fb)
        print("  device-driven %d^3 %.0f MHz: max relres %.2e (recomputed %.2e)" % (nb, fb / 1e6, float(rd.max()),
                                                                                  worst), flush=True)
    for (nb, fb, opb, Bb), X, fl, rr in zip(probs, Xs, fls, rrs):
        assert int(fl.max()) == 0 and float(rr.max()) <= 1e-6, (nb, fb)
        # every column's true residual, one 16-column product
        Bh, Xh = Bb.cpu().numpy(), X.cpu().numpy()
        AX = opb.apply(Xh.T)
        worst = max(_rel(AX[:, q], Bh[q]) for q in range(16))
        assert worst <= 1e-6, (nb, fb, worst)
        # two ports against single solves
        for q in (0, 11):
            xq, flq, relq, _, _ = opb.gmres(Bh[q], restart=50, tol=1e-8, maxit=200)
            assert flq == 0
            assert _rel(Xh[q], xq) <= 1e-5, (nb, fb, q)
        print("  %d^3 %.0f MHz: max relres %.2e (recomputed %.2e)" % (nb, fb / 1e6, float(rr.max()), worst),
              flush=True)
    for p in probs:
        p[2].close()


def _b1plus_setup(n=128):
    import torch
    from marie_amd import adjoint, coil, lib, vie
    dx, f = 0.192 / n, 298e6
    r, eps, sig, _, em = _phantom(n, n, n, dx, f, 0.85, seed=0)
    ports = coil.cylinder_loop_array(8, 0.15, 0.1, 0.05, seg_per_side=8, rings=2, length=0.2)
    Es, Hs = coil.CCL_Excitation(r, em['ko'], 1.0, ports)
    op = lib.Operator(n, n, n)
    op.assemble_N(dx, em['ko'])
    op.assemble_K(dx, em['ko'])
    body = eps > 1.0
    prob = adjoint.JVIEProblem(op, body, Es, f, dx, tol=1e-10, Hinc=Hs)
    nb = prob.idx.size // 3
    dev = torch.device('cuda:0')
    w = torch.tensor(np.exp(-2j * np.pi * np.arange(8) / 4), device=dev)      # CP within each ring
    pts = r.reshape(-1, 3, order='F')[prob.idx[:nb]]
    roi = torch.tensor(np.sum((pts / np.array([0.04, 0.035, 0.05])) ** 2, axis=1) < 1.0, device=dev)

    def objective(e, s, target):
        X = adjoint.jvie_solve(e, s, prob)
        b1 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X, prob), prob), dim=0)
        a = b1.abs()[roi]
        return torch.std(a) / torch.mean(a) + torch.sum((b1 - target).abs() ** 2) / torch.sum(target.abs() ** 2)

    e = torch.tensor(adjoint.to_voxels(eps), device=dev, requires_grad=True)
    s = torch.tensor(adjoint.to_voxels(sig), device=dev, requires_grad=True)
    rng = np.random.default_rng(3)
    with torch.no_grad():
        X0 = adjoint.jvie_solve(e, s, prob)
        b0 = torch.sum(w[:, None] * adjoint.b1_plus(adjoint.total_B(X0, prob), prob), dim=0)
        target = b0 * (1 + 0.2 * torch.tensor(rng.standard_normal(b0.shape), device=dev))
    return prob, body, objective, e, s, target


@pytest.mark.parametrize("which", ["eps_r", "sigma"])
def test_config5_b1plus_gradient_random_voxels(which):
    """Two random body voxels per parameter (drawn among the voxels whose
    gradient magnitude is above the median, so that the check resolves the
    gradient rather than the quotient's noise) against Richardson central
    differences re-solved to 2e-12."""
    import torch
    from marie_amd import adjoint
    print("case B1+ gradient 128^3, 8 ports, random voxels,", which, flush=True)
    prob, body, objective, e, s, target = _b1plus_setup(128)
    J = objective(e, s, target)
    J.backward()
    Jv = float(J.detach())
    grad = (e.grad if which == 'eps_r' else s.grad).cpu().numpy()
    vox = np.nonzero(adjoint.to_voxels(body))[0]
    mag = np.abs(grad[vox])
    cand = vox[mag > np.median(mag)]
    picks = np.random.default_rng(77 if which == 'eps_r' else 78).choice(cand, 2, replace=False)
    prob.tol, prob.maxit = 2e-12, 400
    base = (e if which == 'eps_r' else s).detach()
    for v in picks:
        h = 0.1 * float(base[v])

        def quotient(step):
            vals = []
            for sgn in (1, -1):
                p = base.clone()
                p[v] += sgn * step
                with torch.no_grad():
                    ee, ss = (p, s.detach()) if which == 'eps_r' else (e.detach(), p)
                    vals.append(float(objective(ee, ss, target)))
            return (vals[0] - vals[1]) / (2 * step)

        fd = (4 * quotient(h / 2) - quotient(h)) / 3
        err = abs(fd - grad[v])
        # the re-solves' 2e-12 tolerance leaves J accurate to ~1e-11 relative:
        # a floor of about 6e-11 |J| / h on the Richardson quotient
        floor = 6e-11 * abs(Jv) / h
        print("  voxel %d: autograd %.6e, central difference %.6e, rel diff %.2e (floor %.1e)"
              % (v, grad[v], fd, err / abs(fd), floor / abs(fd)), flush=True)
        assert err <= 1e-5 * abs(fd) + floor, (which, v, fd, grad[v])
        assert abs(fd) > 1e2 * floor
