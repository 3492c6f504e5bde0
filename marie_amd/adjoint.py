"""Differentiable JVIE solve: a torch.autograd.Function over the device solver.

Forward: per-voxel eps_r and sigma (real, the body's voxels) -> the JVIE
form II currents x for each excitation (JVIE_Solver.m:131-156, :242-244,
e_r = eps_r - j sigma/(eo omega) as Scat_Solver.m:144), solved with the
library's FP64 GMRES. Backward: one adjoint solve A^H lambda = g per
excitation (marie_vie_gmres_adjoint_device, mv_AN's 'transp' branch) and the
per-voxel reduction marie_vie_material_gradient_device. The Krylov basis is
not kept between forward and backward: the adjoint solve builds its own
(recomputation instead of storage).

Voxel vectors are flat in the library's voxel order i + L (j + M k)
(MATLAB column-major); `to_voxels` / `from_voxels` convert (L, M, N) arrays.
torch's gradient for a complex output is 2 dJ/dconj(x); the adjoint right-hand
side is therefore grad / 2.
"""
import numpy as np
import torch

from . import lib, vie


def _dense(t):
    """Contiguous memory holding the tensor's values: torch.conj (and views of
    imaginary parts) only set a flag that raw device pointers do not see."""
    return t.resolve_conj().resolve_neg().contiguous()


def to_voxels(a):
    """(L, M, N) array -> flat voxel vector (column-major order)."""
    return np.asarray(a).ravel(order='F')


def from_voxels(v, shape):
    return np.asarray(v).reshape(shape, order='F')


class JVIEProblem:
    """A body (voxel mask), its operator and excitations, at one frequency.

    op: a lib.Operator with fN set (FP64 enabled); body: (L, M, N) bool mask;
    Einc: (L, M, N, 3) complex array or a list of them (one per excitation);
    freq in Hz, dx the voxel size.
    """

    def __init__(self, op, body, Einc, freq, dx, tol=1e-10, restart=50, maxit=200, device='cuda:0', Hinc=None,
                 mode=0, inner_tol=1e-7):
        self.op = op
        self.shape = (op.L, op.M, op.N)
        body = np.asarray(body, bool)
        if body.shape != self.shape:
            raise ValueError('body mask shape %s != operator grid %s' % (body.shape, self.shape))
        n = op.L * op.M * op.N
        idxS = np.nonzero(to_voxels(body))[0]
        self.idx = np.concatenate([idxS, n + idxS, 2 * n + idxS])       # JVIE_Solver.m:147-148
        op.set_index(self.idx)
        self.body = torch.from_numpy(to_voxels(body)).to(device)
        em = vie.EMconstants(freq)
        self.omega_eo = em['omega'] * em['eo']
        self.Gram = dx ** 3
        E = Einc if isinstance(Einc, (list, tuple)) else [Einc]
        self.einc = torch.stack([torch.from_numpy(np.ascontiguousarray(np.asarray(e, np.complex128)
                                                                        .ravel(order='F')[self.idx]))
                                 for e in E]).to(device)
        self.vox = torch.from_numpy(self.idx % n).to(device)
        # GMRES mode of the forward and adjoint solves (0: pgmres; 1-3: complex64
        # products with FP64 refinement, 3 the block-orthogonalised cycles, which
        # needs the operator created with FP32 as well)
        self.tol, self.restart, self.maxit = tol, restart, maxit
        self.mode, self.inner_tol = mode, inner_tol
        self.device = device
        self.hinc = None
        if Hinc is not None:
            Hs = Hinc if isinstance(Hinc, (list, tuple)) else [Hinc]
            self.hinc = torch.stack([torch.from_numpy(np.ascontiguousarray(np.asarray(h, np.complex128)
                                                                            .ravel(order='F')[self.idx]))
                                     for h in Hs]).to(device)

    @property
    def nrhs(self):
        return self.einc.shape[0]

    def material(self, eps_r, sigma):
        """Mc = (e_r - 1)/e_r on the body, 0 in air (form II, Mr = 1)."""
        e_r = torch.complex(eps_r, -sigma / self.omega_eo)
        Mc = (e_r - 1) / e_r
        return torch.where(self.body, Mc, torch.zeros_like(Mc))

    def load_material(self, Mc):
        Mc_h = from_voxels(Mc.detach().cpu().numpy(), self.shape)
        self.op.set_material(np.ones(self.shape), Mc_h, self.Gram)

    def rhs(self, Mc):
        """b = j omega eo Gram Mc Einc on the DOFs (JVIE_Solver.m:151-156 in form II)."""
        return (1j * self.omega_eo * self.Gram) * Mc[self.vox].unsqueeze(0) * self.einc


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, sigma, prob):
        Mc = prob.material(eps_r, sigma)
        prob.load_material(Mc)
        B = _dense(prob.rhs(Mc))
        X = torch.zeros_like(B)
        torch.cuda.synchronize()
        for q in range(prob.nrhs):
            flag, relres, it, rv = prob.op.gmres_device(B[q].data_ptr(), X[q].data_ptr(), prob.restart, prob.tol,
                                                        prob.maxit, mode=prob.mode, inner_tol=prob.inner_tol)
            if flag != 0:
                raise lib.MarieError('forward solve did not converge (relres %.3g)' % relres)
        ctx.prob = prob
        ctx.save_for_backward(X, Mc)
        return X

    @staticmethod
    def backward(ctx, gX):
        X, Mc = ctx.saved_tensors
        prob = ctx.prob
        prob.load_material(Mc)            # another forward may have replaced it
        n = Mc.numel()
        ge = torch.zeros(n, dtype=torch.float64, device=X.device)
        gs = torch.zeros_like(ge)
        te = torch.empty_like(ge)
        ts = torch.empty_like(ge)
        G = _dense(gX / 2)
        for q in range(prob.nrhs):
            lam = torch.zeros_like(G[q])
            torch.cuda.synchronize()
            flag, relres, it, rv = prob.op.gmres_adjoint_device(G[q].data_ptr(), lam.data_ptr(), prob.restart,
                                                                prob.tol, prob.maxit, mode=prob.mode,
                                                                inner_tol=prob.inner_tol)
            if flag != 0:
                raise lib.MarieError('adjoint solve did not converge (relres %.3g)' % relres)
            einc = _dense(prob.einc[q])
            prob.op.material_gradient_device(lam.data_ptr(), _dense(X[q]).data_ptr(), einc.data_ptr(),
                                             prob.omega_eo, te.data_ptr(), ts.data_ptr())
            ge += te
            gs += ts
        return ge, gs, None


class _KApply(torch.autograd.Function):
    """Y = K X on the DOFs (mv_K.m); K is symmetric as a matrix (each
    component of its kernel is odd in its own axis and enters through an
    antisymmetric cross product), so the backward product is
    K^H g = conj(K conj(g))."""

    @staticmethod
    def forward(ctx, X, prob):
        ctx.prob = prob
        return _k_product(prob, X)

    @staticmethod
    def backward(ctx, g):
        return torch.conj(_k_product(ctx.prob, torch.conj(g))), None


def _k_product(prob, X):
    # the product runs on the handle's stream: torch's work on X must be done
    # before it starts, and Y must be complete before torch reads it
    Xc = _dense(X)
    Y = torch.empty_like(Xc)
    torch.cuda.synchronize()
    prob.op.apply_device(Xc.data_ptr(), Y.data_ptr(), lib.OP_K, False, lib.PREC_F64, Xc.shape[0])
    torch.cuda.synchronize()
    return Y


def k_apply(X, prob):
    """K X (nrhs x nidx) on the problem's DOFs; the handle must hold fK
    (assemble_K or set_fK)."""
    return _KApply.apply(X, prob)


def total_B(X, prob):
    """B = mu0 (Hinc + K J / Gram) on the DOFs (H_field_Kop.m:69-76,
    Scat_Solver.m:215-216), one row per excitation; needs Hinc in the problem."""
    if prob.hinc is None:
        raise ValueError('JVIEProblem was built without Hinc')
    return 4e-7 * np.pi * (prob.hinc + k_apply(X, prob) / prob.Gram)


def b1_plus(B, prob):
    """(Bx + j By)/2 per body voxel (the usual MR definition; the reference defines
    none), from B on the DOFs: nrhs x n_body."""
    nb = prob.idx.size // 3
    return (B[:, :nb] + 1j * B[:, nb:2 * nb]) / 2


def jvie_solve(eps_r, sigma, prob):
    """x (nrhs x nidx complex128) for flat voxel eps_r, sigma (float64 CUDA tensors)."""
    return _Solve.apply(eps_r, sigma, prob)


def interior_field(X, eps_r, sigma, prob):
    """E = J / tau on the body DOFs, tau = j omega eo (e_r - 1) (JVIE_Solver.m:151)."""
    e_r = torch.complex(eps_r, -sigma / prob.omega_eo)
    tau = 1j * prob.omega_eo * (e_r - 1)
    return X / tau[prob.vox].unsqueeze(0)
