// Per-voxel gradient of a real objective with respect to the body's relative
// permittivity and conductivity, from the state x and the adjoint state.
//
// JVIE form II (JVIE_Solver.m:131-156, :242-244) with e_r = eps_r - j sigma /
// (eo omega) (Scat_Solver.m:144): Mr = 1, Mc = (e_r - 1)/e_r and
// b = j omega eo Gram Mc Einc, so the residual is
//     F(x, Mc) = A x - b = Gram x - Mc (N x + j omega eo Gram Einc) = Gram x - Mc u.
// Mc depends on e_r through dMc/de_r = 1/e_r^2 = (1 - Mc)^2, and for a real
// objective J(x) with g = dJ/dconj(x), the adjoint state lambda solving
// A^H lambda = g (mv_AN.m:65-98, the 'transp' branch) gives
//     S_v = sum_c conj(lambda_vc) (1 - Mc_v)^2 u_vc
//     dJ/d eps_r(v) = 2 Re S_v,   dJ/d sigma(v) = 2 Im S_v / (omega eo).
// Only voxels of the DOF index carry a gradient (air voxels are outside the
// unknowns); others get 0.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

namespace marie {
namespace {

// u = u + c * einc (c = j omega eo Gram)
__global__ void k_add_scaled(c2<double>* __restrict__ u, const c2<double>* __restrict__ einc, double cre, double cim,
                             int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const c2<double> a = einc[e];
        u[e] = c2<double>{u[e].x + (cre * a.x - cim * a.y), u[e].y + (cre * a.y + cim * a.x)};
    }
}

__global__ void k_material_grad(const c2<double>* __restrict__ lam, const c2<double>* __restrict__ u,
                                const int* __restrict__ pos, const c2<double>* __restrict__ Mc, int64_t nv,
                                double inv_we, double* __restrict__ ge, double* __restrict__ gs) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        double sx = 0.0, sy = 0.0;
        for (int c = 0; c < 3; ++c) {
            const int p = pos[c * nv + v];
            if (p < 0) continue;
            const c2<double> l = lam[p], w = u[p];
            // conj(l) * w
            sx += l.x * w.x + l.y * w.y;
            sy += l.x * w.y - l.y * w.x;
        }
        // times (1 - Mc)^2
        const c2<double> m = Mc[v];
        const double ax = 1.0 - m.x, ay = -m.y;
        const double fx = ax * ax - ay * ay, fy = 2.0 * ax * ay;
        const double tx = sx * fx - sy * fy, ty = sx * fy + sy * fx;
        ge[v] = 2.0 * tx;
        gs[v] = 2.0 * ty * inv_we;
    }
}

}  // namespace
}  // namespace marie

using namespace marie;

extern "C" int marie_vie_material_gradient_device(marie_vie_op* hh, const void* lambda, const void* x,
                                                  const void* einc, double omega_eo, double* grad_eps,
                                                  double* grad_sigma) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!lambda || !x || !einc || !grad_eps || !grad_sigma || !(omega_eo > 0.0)) {
        set_error("invalid material gradient arguments");
        return MARIE_ERR_ARG;
    }
    if (!h->have_fN || !h->have_mat || !h->have_idx || !(h->precisions & PREC_F64) || !h->Mc64 || !h->pos) {
        set_error("operator, index or FP64 material not set");
        return MARIE_ERR_STATE;
    }
    const int64_t nv = h->nvox();
    const hipStream_t s = h->stream;
    c2<double>* u = nullptr;
    if (h->nidx > 0) MARIE_HIP(hipMalloc(&u, h->nidx * sizeof(c2<double>)));
    int st = OK;
    if (h->nidx > 0) st = apply(h, OP_N, 0, PREC_F64, x, u, 1, s);
    if (st == OK && h->nidx > 0) {
        const int64_t g = (h->nidx + 255) / 256 < 4096 ? (h->nidx + 255) / 256 : 4096;
        hipLaunchKernelGGL(k_add_scaled, dim3((unsigned)g), dim3(256), 0, s, u, (const c2<double>*)einc, 0.0,
                           omega_eo * h->Gram, h->nidx);
    }
    if (st == OK) {
        const int64_t g = (nv + 255) / 256 < 4096 ? (nv + 255) / 256 : 4096;
        hipLaunchKernelGGL(k_material_grad, dim3((unsigned)g), dim3(256), 0, s, (const c2<double>*)lambda, u, h->pos,
                           (const c2<double>*)h->Mc64, nv, 1.0 / omega_eo, grad_eps, grad_sigma);
        if (hipGetLastError() != hipSuccess) {
            set_error("material gradient launch failed");
            st = ERR_HIP;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess && st == OK) {
        set_error("material gradient failed on the device");
        st = ERR_HIP;
    }
    if (u) (void)hipFree(u);
    return st;
}
