"""GPU parity of the remaining JVIE_Solver branches: the preconditioner
operator (mv_AN with 1./Mr, 1./Mr - 1), left/right-preconditioned GMRES
(pgmres L1 / R1) and BiCGStab (plain and preconditioned), each against the
oracle on the same seeded problems; forms I and II."""
import numpy as np
import pytest

from oracle import fields, krylov, mvp
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _problem(n=12, dx=6.25e-3, f=298.2e6, form=2, eps=50.0, sig=0.5):
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    inside = X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2
    em = fields.em_constants(f)
    e_r = np.where(inside, eps - 1j * sig / (em['eo'] * em['omega']), 1.0 + 0j)
    Einc, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=form)
    return fN, Mr, Mc, Gram, idx, b


def _op(fN, Mr, Mc, Gram, idx):
    from marie_amd import lib
    n = Mr.shape[0]
    op = lib.Operator(n, n, n, precisions=lib.PREC_F64 | lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    return op


@pytest.mark.parametrize("form", [1, 2])
def test_preconditioner_operator(form):
    """OP_P = mv_AN(., 1./Mr, 1./Mr - 1): the full product in form I, and
    exactly Gram*x in form II (Mr = 1), both directions."""
    from marie_amd import lib
    print("case OP_P form", form, flush=True)
    fN, Mr, Mc, Gram, idx, b = _problem(form=form)
    op = _op(fN, Mr, Mc, Gram, idx)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((idx.size, 2)) + 1j * rng.standard_normal((idx.size, 2))
    for transp in ('notransp', 'transp'):
        y = op.apply(x, lib.OP_P, transp=transp == 'transp')
        for q in range(2):
            yr = mvp.mv_AN(x[:, q], fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, transp, idx)
            assert _rel(y[:, q], yr) <= 1e-12
            if form == 2:
                assert np.array_equal(y[:, q], Gram * x[:, q])
    # the operator follows a material change (derived data recomputed)
    op.set_material(2.0 * Mr, Mc, Gram)
    y = op.apply(x[:, 0], lib.OP_P)
    yr = mvp.mv_AN(x[:, 0], fN, 0.5 / Mr, 0.5 / Mr - 1.0, Gram, 'notransp', idx)
    assert _rel(y, yr) <= 1e-12


def test_preconditioner_operator_complex64_only_handle():
    """OP_P on a handle created with complex64 only (no FP64 material on the
    device): the preconditioner's material is derived from the complex64 Mr,
    and the product matches the oracle's FP64 mv_AN(., 1./Mr, 1./Mr - 1) to
    complex64 accuracy, in both directions."""
    from marie_amd import lib
    print("case OP_P complex64-only handle", flush=True)
    fN, Mr, Mc, Gram, idx, b = _problem(form=1)
    n = Mr.shape[0]
    op = lib.Operator(n, n, n, precisions=lib.PREC_F32)
    op.set_fN(fN)
    op.set_material(Mr, Mc, Gram)
    op.set_index(idx)
    rng = np.random.default_rng(6)
    x = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    for transp in ('notransp', 'transp'):
        y = op.apply(x, lib.OP_P, transp=transp == 'transp')
        yr = mvp.mv_AN(x, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, transp, idx)
        assert _rel(y, yr) <= 2e-6, transp


@pytest.mark.parametrize("side,form", [('R', 1), ('L', 1), ('R', 2), ('L', 2)])
def test_preconditioned_gmres_parity(side, form):
    from marie_amd import lib
    print("case pgmres", side, "form", form, flush=True)
    fN, Mr, Mc, Gram, idx, b = _problem(form=form)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    P = lambda v: mvp.mv_AN(v, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, 'notransp', idx)
    if side == 'R':
        xr, flr, rr, itr, rvr = krylov.pgmres(A, b, 20, 1e-8, 20, None, None, P)
    else:
        xr, flr, rr, itr, rvr = krylov.pgmres(A, b, 20, 1e-8, 20, P)
    op = _op(fN, Mr, Mc, Gram, idx)
    s = lib.PRECOND_RIGHT if side == 'R' else lib.PRECOND_LEFT
    x, fl, rel, it, rv = op.gmres_precond(b, restart=20, tol=1e-8, maxit=20, side=s)
    assert fl == flr == 0
    assert it == itr and rv.size == rvr.size
    assert np.max(np.abs(rv - rvr)) <= 1e-9
    assert _rel(x, xr) <= 1e-8
    assert abs(rel - rr) <= 1e-9


@pytest.mark.parametrize("precond,form", [(False, 2), (True, 1), (True, 2), (False, 1)])
def test_bicgstab_parity(precond, form):
    """Same recurrences as the oracle: the same number of half steps, resvec
    to rounding until the iterates separate, the solution to the tolerance."""
    print("case bicgstab precond", precond, "form", form, flush=True)
    fN, Mr, Mc, Gram, idx, b = _problem(form=form)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    M = (lambda v: mvp.mv_AN(v, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, 'notransp', idx)) if precond else None
    tol = 1e-9
    xr, flr, rr, itr, rvr = krylov.bicgstab(A, b, tol, 400, M=M)
    op = _op(fN, Mr, Mc, Gram, idx)
    x, fl, rel, it, rv = op.bicgstab(b, tol=tol, maxit=400, precond=precond)
    assert flr == 0 and fl == 0 and rel <= tol
    # a short-recurrence method: rounding differences grow (a perturbation of
    # b by 2^-50 moves the oracle's own count by up to 2.5 steps here), so the
    # counts may differ by a few steps; the first steps agree to rounding
    assert abs(it - itr) <= 4.0
    k = min(rv.size, rvr.size, 12)
    assert np.max(np.abs(rv[:k] - rvr[:k]) / rvr[0]) <= 1e-10
    assert _rel(x, xr) <= 100 * tol


def test_bicgstab_fixed_iterations_and_edges():
    """maxit reached: flag 1, exactly 2 maxit + 1 resvec entries and the oracle's
    iterate after those steps; x0 given; zero right-hand side; a converged
    initial guess takes no step."""
    print("case bicgstab maxit / x0 / zero", flush=True)
    fN, Mr, Mc, Gram, idx, b = _problem(n=10)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    op = _op(fN, Mr, Mc, Gram, idx)
    rng = np.random.default_rng(2)
    x0 = 1e-3 * (rng.standard_normal(b.size) + 1j * rng.standard_normal(b.size)) * np.abs(b).max()
    for k in (1, 3, 6):
        xr, flr, rr, itr, rvr = krylov.bicgstab(A, b, 1e-14, k, x0=x0)
        x, fl, rel, it, rv = op.bicgstab(b, x0=x0, tol=1e-14, maxit=k)
        assert fl == flr == 1 and it == itr == k
        assert rv.size == rvr.size == 2 * k + 1
        assert np.max(np.abs(rv - rvr) / rvr[0]) <= 1e-11
        assert _rel(x, xr) <= 1e-11
        assert abs(rel - rr) <= 1e-11
    x, fl, rel, it, rv = op.bicgstab(np.zeros_like(b))
    assert fl == 0 and it == 0 and np.all(x == 0)
    xs, fl, rel, it, rv = op.bicgstab(b, tol=1e-10, maxit=300)
    x, fl2, rel2, it2, rv2 = op.bicgstab(b, x0=xs, tol=1e-6, maxit=300)
    assert fl2 == 0 and it2 == 0 and rv2.size == 1 and np.array_equal(x, xs)


def test_jvie_solver_dispatch(tmp_path):
    """JVIE_Solver's default (BiCGStab), 'G' and the preconditioned branches
    reach the same currents on a small sphere in form I; each run appends the
    reference's summary (JVIE_Solver.m:159-166, :335) to one log file, whose
    iteration counts and residues are read back."""
    import re
    from marie_amd import vie
    print("case JVIE_Solver dispatch", flush=True)
    n, dx, f = 12, 6.25e-3, 298.2e6
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    em = vie.EMconstants(f)
    e_r = np.where(X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2, 50 - 1j * 0.5 / (em['eo'] * em['omega']), 1.0)
    Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    ref, flag, *_ = vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-10, 1, 'G')
    assert flag == 0
    log = tmp_path / 'vie.log'
    runs = []
    for solver, precond in (('B', 0), ('M', 'L'), ('G', 'R'), ('G', 'L'), ('B', 'R'), ('D', 0), ('D', 'R')):
        J, flag, relres, it, resvec, t = vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-8, 1, solver, precond,
                                                         logfile=str(log))
        assert flag == 0 and relres <= 1e-8, (solver, precond)
        assert _rel(J, ref) <= 1e-6, (solver, precond)
        runs.append((solver, precond, len(resvec), relres))
    text = log.read_text()
    assert text.count(' Domain:                %dx%dx%d voxels' % (n, n, n)) == len(runs)
    assert text.count(' # DOFS:                %d' % (3 * n ** 3)) == len(runs)
    heads = re.findall(r'\n GPU (GMRES|BICGSTAB) solver[^\n]*?, (wo?/ [A-Za-z ]*preconditioner)', text)
    times = re.findall(r'Time  = [0-9.]+ \[sec\] for (\d+) iterations, flag (\d+), residue \S+, '
                       r'relative residue (\S+) ', text)
    assert len(heads) == len(times) == len(runs)
    for (solver, precond, nres, relres), (kind, pc), (n_it, fl, rr) in zip(runs, heads, times):
        assert kind == ('GMRES' if solver in 'GD' else 'BICGSTAB')
        assert pc.startswith('wo/') == (precond not in ('L', 'R'))
        assert int(n_it) == nres and int(fl) == 0 and abs(float(rr) - relres) <= 1e-5 * relres
    assert text.count('VIE Solve finished.') == len(runs)


@pytest.mark.parametrize("precision", ['fp32-refined', 'fp32-refined-c64', 'fp32-refined-blocks'])
def test_jvie_solver_refined_precisions(precision):
    """JVIE_Solver's complex64-refined GMRES variants (modes 1, 2 and 3) in
    form II reach the FP64 tolerance on the true residual and the FP64
    currents within that solve's own error."""
    from marie_amd import vie
    print("case JVIE_Solver", precision, flush=True)
    n, dx, f = 12, 6.25e-3, 298.2e6
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    em = vie.EMconstants(f)
    e_r = np.where(X ** 2 + Y ** 2 + Z ** 2 < (0.45 * n * dx) ** 2, 50 - 1j * 0.5 / (em['eo'] * em['omega']), 1.0)
    Einc, _ = vie.PlaneWave_Excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    ref, flag, *_ = vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-11, 2, 'G')
    J64, flag64, rel64, *_ = vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-6, 2, 'G')
    J, flag, relres, it, resvec, t = vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-6, 2, 'G', inner_it=40,
                                                     precision=precision, inner_tol=1e-7)
    assert flag == 0 and flag64 == 0 and relres <= 1e-6
    assert _rel(J, ref) <= max(5 * _rel(J64, ref), 1e-6)
    with pytest.raises(ValueError):
        vie.JVIE_Solver(Einc, e_r, r, f, fN, 1e-6, 2, 'G', precision='fp16')


def _hetero_problem(n=12, dx=6.25e-3, f=298.2e6, form=2, seed=3):
    """Two-shell sphere with seeded permittivities (forms I and II differ)."""
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    r = np.stack([X, Y, Z], axis=-1)
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rng = np.random.default_rng(seed)
    em = fields.em_constants(f)
    eps = np.where(R < 0.25 * n * dx, 20 + 60 * rng.random(), np.where(R < 0.45 * n * dx, 20 + 60 * rng.random(), 1.0))
    sig = np.where(R < 0.45 * n * dx, 0.1 + 1.9 * rng.random(R.shape), 0.0)
    e_r = eps - 1j * sig / (em['eo'] * em['omega'])
    Einc, _ = fields.plane_wave_excitation(r, [0, 0, em['ko']], em['omega_mu'], [1, 0, 0])
    fN = oop.get_operators(n, n, n, dx, f, 'N')
    Mr, Mc, Gram, idx, b = fields.jvie_setup(Einc, e_r, dx, f, form=form)
    return fN, Mr, Mc, Gram, idx, b


@pytest.mark.parametrize("side,form", [(None, 2), ('R', 1), ('L', 1), (None, 1)])
def test_gmres_dr_parity(side, form):
    """GMRES with deflated restarts against the oracle's pgmresDR: same cycle
    lengths, resvec to 1e-8 (the harmonic Ritz eigenproblems are solved by
    different codes), solution to the tolerance."""
    from marie_amd import lib
    print("case gmres_dr", side, "form", form, flush=True)
    fN, Mr, Mc, Gram, idx, b = _hetero_problem(form=form)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    P = lambda v: mvp.mv_AN(v, fN, 1.0 / Mr, 1.0 / Mr - 1.0, Gram, 'notransp', idx)
    kw = {} if side is None else ({'R1': P} if side == 'R' else {'L1': P})
    xr, flr, rr, itr, rvr = krylov.pgmresDR(A, b, 10, 1e-8, 60, 4, **kw)
    op = _op(fN, Mr, Mc, Gram, idx)
    s = {None: lib.PRECOND_NONE, 'R': lib.PRECOND_RIGHT, 'L': lib.PRECOND_LEFT}[side]
    x, fl, rel, it, rv = op.gmres_dr(b, restart=10, tol=1e-8, maxit=60, ritz=4, side=s)
    assert flr == 0 and fl == 0 and rel <= 1e-8
    assert rv.size == rvr.size and it == itr
    assert np.allclose(rv[:11], rvr[:11], rtol=1e-10, atol=0)        # the first cycle is GMRES(10)
    assert np.max(np.abs(rv - rvr)) <= 1e-8
    assert _rel(x, xr) <= 1e-7


def test_gmres_dr_arguments():
    """ritz beyond restart - 1 is limited to it; ritz < 1 takes the default rule;
    a first cycle that converges early needs no recycled block; zero rhs."""
    from marie_amd import lib
    print("case gmres_dr arguments", flush=True)
    fN, Mr, Mc, Gram, idx, b = _hetero_problem(n=10)
    A = lambda v: mvp.mv_AN(v, fN, Mr, Mc, Gram, 'notransp', idx)
    op = _op(fN, Mr, Mc, Gram, idx)
    for ritz, ref_ritz in ((9, 5), (0, None)):
        xr, flr, rr, itr, rvr = krylov.pgmresDR(A, b, 6, 1e-7, 60, ref_ritz)
        x, fl, rel, it, rv = op.gmres_dr(b, restart=6, tol=1e-7, maxit=60, ritz=ritz)
        assert fl == flr == 0 and rv.size == rvr.size and np.max(np.abs(rv - rvr)) <= 1e-8
    x, fl, rel, it, rv = op.gmres_dr(b, restart=60, tol=1e-3, maxit=5, ritz=5)
    assert fl == 0 and it[1] == 0
    x, fl, rel, it, rv = op.gmres_dr(np.zeros_like(b))
    assert fl == 0 and np.all(x == 0)
    with pytest.raises(lib.MarieError):
        op.gmres_dr(b, restart=1)
