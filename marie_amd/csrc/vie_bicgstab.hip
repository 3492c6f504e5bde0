// Device-resident BiCGStab, the solver JVIE_Solver.m:463-490 calls
// (MATLAB's bicgstab(afun, Vrhs, tol, max_it[, pfun]) on mv_AN, with the
// preconditioner handle fPGPU = mv_AN(., 1./Mr, 1./Mr - 1) as M).
//
// Recurrences (van der Vorst 1992, shadow residual r~ = r0, preconditioned
// vectors p^ = M(p), s^ = M(s) as MATLAB applies a function-handle M):
//   rho = <r~, r>;  p = r + beta (p - omega v),  beta = (rho/rho_old)(alpha/omega)
//   v = A p^;  alpha = rho / <r~, v>;  s = r - alpha v;  x += alpha p^
//   t = A s^;  omega = <t, s>/<t, t>;  x += omega s^;  r = s - omega t
// resvec holds the absolute residual norm of every half step (||r0||, ||s||,
// ||r||, ...); the iteration stops when a half- or full-step norm is
// <= tol*||b||, iter counts half steps / 2; a zero <r~,r>, <r~,v>, <t,t> or
// omega ends it as a breakdown (flag 4 unless the true residual meets tol);
// flag and relres come from the true residual ||b - A x||/||b||.
//
// Every scalar stays on the device: dot products reduce in fixed order into
// a small state record that the update kernels read, and a kernel that finds
// the iteration finished sets `done`, after which every later update is a
// no-op. The host queues iteration k+1 before it reads iteration k's state
// from pinned memory, so the device never waits on the host; after the stop
// the queued iteration changes nothing (its products write only v, t and the
// preconditioned vectors, which the result never reads).
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"
#include "vie_reduce.hpp"

#include <cmath>
#include <vector>

namespace marie {
namespace {

constexpr int BT = 256;     // threads per block
constexpr int BG = 1024;    // max blocks (partial sums per reduction)

struct BiState {
    c2<double> rho, rho_old, alpha, omega, beta;
    double tolb;            // tol * ||b||
    int done;               // 0 running, 1 converged, 2 breakdown
    int half;               // completed half steps
    int nres;               // resvec entries written
    int iter;               // iterations started
};

__device__ __forceinline__ c2<double> cdiv(c2<double> a, c2<double> b) {
    const double d = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ __forceinline__ c2<double> cm(c2<double> a, c2<double> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Partial sums: part[q * G + block] for q < W (W real quantities)
template <int W>
__device__ __forceinline__ void store_partials(double (&acc)[W], double* part, int G) {
    __shared__ double sh[BT / 64 * W];
    block_reduce<W, BT>(acc, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < W; ++q) part[q * G + blockIdx.x] = acc[q];
}

// thread 0 gets sum_b part[q * G + b], fixed order
template <int W>
__device__ __forceinline__ void sum_partials(const double* part, int G, double (&acc)[W]) {
    __shared__ double sh[BT / 64 * W];
#pragma unroll
    for (int q = 0; q < W; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < G; b += BT)
#pragma unroll
        for (int q = 0; q < W; ++q) acc[q] += part[q * G + b];
    block_reduce<W, BT>(acc, sh);
}

// ---- start of iteration: breakdown test on rho, beta; then p
__global__ void k_sc_beta(BiState* st) {
    if (threadIdx.x != 0 || st->done) return;
    st->iter += 1;
    if (st->rho.x == 0.0 && st->rho.y == 0.0) {
        st->done = 2;
        return;
    }
    if (st->iter > 1) st->beta = cm(cdiv(st->rho, st->rho_old), cdiv(st->alpha, st->omega));
}

__global__ __launch_bounds__(BT) void k_update_p(const BiState* __restrict__ st, c2<double>* __restrict__ p,
                                                 const c2<double>* __restrict__ r, const c2<double>* __restrict__ v,
                                                 int64_t n) {
    if (st->done) return;
    const bool first = st->iter == 1;
    const c2<double> beta = st->beta, om = st->omega;
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)gridDim.x * BT) {
        const c2<double> re = r[e];
        if (first) {
            p[e] = re;
        } else {
            const c2<double> ve = v[e], pe = p[e];
            const c2<double> ov = cm(om, ve);
            const c2<double> d = {pe.x - ov.x, pe.y - ov.y};
            const c2<double> bd = cm(beta, d);
            p[e] = c2<double>{re.x + bd.x, re.y + bd.y};
        }
    }
}

// ---- <r~, v> -> alpha
__global__ __launch_bounds__(BT) void k_dot_rtv(const BiState* __restrict__ st, const c2<double>* __restrict__ rt,
                                                const c2<double>* __restrict__ v, int64_t n, double* part, int G) {
    if (st->done) return;
    double acc[2] = {0.0, 0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> a = rt[e], b = v[e];
        acc[0] += a.x * b.x + a.y * b.y;
        acc[1] += a.x * b.y - a.y * b.x;
    }
    store_partials<2>(acc, part, G);
}

__global__ __launch_bounds__(BT) void k_sc_alpha(BiState* st, const double* part, int G) {
    if (st->done) return;
    double acc[2];
    sum_partials<2>(part, G, acc);
    if (threadIdx.x != 0) return;
    if (acc[0] == 0.0 && acc[1] == 0.0) {
        st->done = 2;
        return;
    }
    st->alpha = cdiv(st->rho, c2<double>{acc[0], acc[1]});
}

// ---- s = r - alpha v, x += alpha p^, ||s||^2
__global__ __launch_bounds__(BT) void k_update_s(const BiState* __restrict__ st, c2<double>* __restrict__ s,
                                                 const c2<double>* __restrict__ r, const c2<double>* __restrict__ v,
                                                 c2<double>* __restrict__ x, const c2<double>* __restrict__ ph,
                                                 int64_t n, double* part, int G) {
    if (st->done) return;
    const c2<double> al = st->alpha;
    double acc[1] = {0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> av = cm(al, v[e]), re = r[e];
        const c2<double> se = {re.x - av.x, re.y - av.y};
        s[e] = se;
        const c2<double> ap = cm(al, ph[e]), xe = x[e];
        x[e] = c2<double>{xe.x + ap.x, xe.y + ap.y};
        acc[0] += se.x * se.x + se.y * se.y;
    }
    store_partials<1>(acc, part, G);
}

// residual norm of a half step: resvec entry, convergence test
__global__ __launch_bounds__(BT) void k_sc_res(BiState* st, const double* part, int G, double* res) {
    if (st->done) return;
    double acc[1];
    sum_partials<1>(part, G, acc);
    if (threadIdx.x != 0) return;
    const double nr = sqrt(acc[0]);
    res[st->nres] = nr;
    st->nres += 1;
    st->half += 1;
    if (nr <= st->tolb) st->done = 1;
}

// ---- <t, s>, <t, t> -> omega
__global__ __launch_bounds__(BT) void k_dot_ts(const BiState* __restrict__ st, const c2<double>* __restrict__ t,
                                               const c2<double>* __restrict__ s, int64_t n, double* part, int G) {
    if (st->done) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> a = t[e], b = s[e];
        acc[0] += a.x * b.x + a.y * b.y;
        acc[1] += a.x * b.y - a.y * b.x;
        acc[2] += a.x * a.x + a.y * a.y;
    }
    store_partials<3>(acc, part, G);
}

__global__ __launch_bounds__(BT) void k_sc_omega(BiState* st, const double* part, int G) {
    if (st->done) return;
    double acc[3];
    sum_partials<3>(part, G, acc);
    if (threadIdx.x != 0) return;
    if (acc[2] == 0.0) {
        st->done = 2;
        return;
    }
    const c2<double> om = {acc[0] / acc[2], acc[1] / acc[2]};
    if (om.x == 0.0 && om.y == 0.0) {
        st->done = 2;
        return;
    }
    st->omega = om;
}

// ---- x += omega s^, r = s - omega t, ||r||^2 and <r~, r>
__global__ __launch_bounds__(BT) void k_update_r(const BiState* __restrict__ st, c2<double>* __restrict__ x,
                                                 const c2<double>* __restrict__ sh, c2<double>* __restrict__ r,
                                                 const c2<double>* __restrict__ s, const c2<double>* __restrict__ t,
                                                 const c2<double>* __restrict__ rt, int64_t n, double* part, int G) {
    if (st->done) return;
    const c2<double> om = st->omega;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> os = cm(om, sh[e]), xe = x[e];
        x[e] = c2<double>{xe.x + os.x, xe.y + os.y};
        const c2<double> ot = cm(om, t[e]), se = s[e];
        const c2<double> re = {se.x - ot.x, se.y - ot.y};
        r[e] = re;
        const c2<double> a = rt[e];
        acc[0] += re.x * re.x + re.y * re.y;
        acc[1] += a.x * re.x + a.y * re.y;
        acc[2] += a.x * re.y - a.y * re.x;
    }
    store_partials<3>(acc, part, G);
}

__global__ __launch_bounds__(BT) void k_sc_r(BiState* st, const double* part, int G, double* res) {
    if (st->done) return;
    double acc[3];
    sum_partials<3>(part, G, acc);
    if (threadIdx.x != 0) return;
    const double nr = sqrt(acc[0]);
    res[st->nres] = nr;
    st->nres += 1;
    st->half += 1;
    if (nr <= st->tolb) {
        st->done = 1;
        return;
    }
    st->rho_old = st->rho;
    st->rho = c2<double>{acc[1], acc[2]};
}

// ---- set-up: r = b - A x0 (or b), r~ = r, ||b||, ||r0||, rho = <r~, r>
__global__ __launch_bounds__(BT) void k_setup_partials(const c2<double>* __restrict__ b,
                                                       const c2<double>* __restrict__ r, int64_t n, double* part,
                                                       int G) {
    double acc[2] = {0.0, 0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> be = b[e], re = r[e];
        acc[0] += be.x * be.x + be.y * be.y;
        acc[1] += re.x * re.x + re.y * re.y;
    }
    store_partials<2>(acc, part, G);
}

__global__ __launch_bounds__(BT) void k_sc_setup(BiState* st, const double* part, int G, double tol, double* res,
                                                 double* bnorm) {
    double acc[2];
    sum_partials<2>(part, G, acc);
    if (threadIdx.x != 0) return;
    const double nb = sqrt(acc[0]), nr = sqrt(acc[1]);
    *bnorm = nb;
    st->tolb = tol * nb;
    st->done = (nr <= st->tolb) ? 1 : 0;
    st->half = 0;
    st->iter = 0;
    st->nres = 1;
    res[0] = nr;
    st->rho = c2<double>{acc[1], 0.0};     // <r~, r> with r~ = r
    st->rho_old = st->alpha = st->omega = st->beta = c2<double>{1.0, 0.0};
}

// r = b - y
__global__ void k_sub(c2<double>* __restrict__ r, const c2<double>* __restrict__ b, const c2<double>* __restrict__ y,
                      int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const c2<double> be = b[e], ye = y[e];
        r[e] = c2<double>{be.x - ye.x, be.y - ye.y};
    }
}

__global__ __launch_bounds__(BT) void k_norm_partials(const c2<double>* __restrict__ v, int64_t n, double* part,
                                                      int G) {
    double acc[1] = {0.0};
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT) {
        const c2<double> a = v[e];
        acc[0] += a.x * a.x + a.y * a.y;
    }
    store_partials<1>(acc, part, G);
}

__global__ __launch_bounds__(BT) void k_sum1(const double* part, int G, double* out) {
    double acc[1];
    sum_partials<1>(part, G, acc);
    if (threadIdx.x == 0) *out = acc[0];
}

struct BiWs {
    c2<double>*r = nullptr, *rt = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr, *ph = nullptr,
    *sh = nullptr;
    double* part = nullptr;        // 3 x BG
    double* res = nullptr;         // 2 maxit + 1
    double* scal = nullptr;        // 2: ||b||, scratch
    BiState* st = nullptr;
    BiState* hst = nullptr;        // pinned: 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    void free_all() {
        void* q[] = {r, rt, p, v, s, t, ph, sh, part, res, scal, st};
        for (void* a : q)
            if (a) (void)hipFree(a);
        if (hst) (void)hipHostFree(hst);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

int grid_for(int64_t n) {
    int64_t g = (n + BT * 4 - 1) / (BT * 4);
    if (g < 1) g = 1;
    if (g > BG) g = BG;
    return (int)g;
}

}  // namespace

int bicgstab(Op* h, const c2<double>* b, c2<double>* x, double tol, int maxit, int precond, int* flag,
             double* relres, double* iter, std::vector<double>& resvec) {
    const int64_t n = h->nidx;
    const hipStream_t s = h->stream;
    const int G = grid_for(n);
    BiWs w;
    auto done = [&](int code) {
        (void)hipStreamSynchronize(s);
        w.free_all();
        return code;
    };
    auto A = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the BiCGStab workspace");
            return ERR_ALLOC;
        }
        return OK;
    };
    const size_t vb = (size_t)n * sizeof(c2<double>);
    int st = A((void**)&w.r, vb);
    if (!st) st = A((void**)&w.rt, vb);
    if (!st) st = A((void**)&w.p, vb);
    if (!st) st = A((void**)&w.v, vb);
    if (!st) st = A((void**)&w.s, vb);
    if (!st) st = A((void**)&w.t, vb);
    if (!st && precond) st = A((void**)&w.ph, vb);
    if (!st && precond) st = A((void**)&w.sh, vb);
    if (!st) st = A((void**)&w.part, 3 * BG * sizeof(double));
    if (!st) st = A((void**)&w.res, (size_t)(2 * maxit + 1) * sizeof(double));
    if (!st) st = A((void**)&w.scal, 2 * sizeof(double));
    if (!st) st = A((void**)&w.st, sizeof(BiState));
    if (!st && hipHostMalloc((void**)&w.hst, 2 * sizeof(BiState), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        w.hst = nullptr;
        set_error("pinned host allocation failed");
        st = ERR_ALLOC;
    }
    for (int e = 0; e < 2 && !st; ++e)
        if (hipEventCreateWithFlags(&w.ev[e], hipEventDisableTiming) != hipSuccess) {
            set_error("hipEventCreate failed");
            st = ERR_HIP;
        }
    if (st) return done(st);
    c2<double>* ph = precond ? w.ph : w.p;
    c2<double>* sh = precond ? w.sh : w.s;
    const int eg = G;

    // ||x0|| decides between r = b - A x0 and r = b
    hipLaunchKernelGGL(k_norm_partials, dim3(G), dim3(BT), 0, s, x, n, w.part, G);
    hipLaunchKernelGGL(k_sum1, dim3(1), dim3(BT), 0, s, w.part, G, w.scal + 1);
    MARIE_HIP(hipGetLastError());
    double xn = 0.0, bnorm = 0.0;
    MARIE_HIP(hipMemcpyAsync(&xn, w.scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipStreamSynchronize(s));
    if (xn > 0.0) {
        st = apply(h, OP_AN, 0, PREC_F64, x, w.t, 1, s);
        if (st) return done(st);
        hipLaunchKernelGGL(k_sub, dim3(eg), dim3(BT), 0, s, w.r, b, w.t, n);
    } else {
        MARIE_HIP(hipMemcpyAsync(w.r, b, vb, hipMemcpyDeviceToDevice, s));
    }
    MARIE_HIP(hipMemcpyAsync(w.rt, w.r, vb, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_setup_partials, dim3(G), dim3(BT), 0, s, b, w.r, n, w.part, G);
    hipLaunchKernelGGL(k_sc_setup, dim3(1), dim3(BT), 0, s, w.st, w.part, G, tol, w.res, w.scal);
    MARIE_HIP(hipGetLastError());
    MARIE_HIP(hipMemcpyAsync(&bnorm, w.scal, sizeof(double), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipMemcpyAsync(w.hst, w.st, sizeof(BiState), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipStreamSynchronize(s));
    if (bnorm == 0.0) {
        // zero right-hand side: x = 0 (as the GMRES entry points)
        MARIE_HIP(hipMemsetAsync(x, 0, vb, s));
        *flag = 0;
        *relres = 0.0;
        *iter = 0.0;
        resvec.assign(1, 0.0);
        return done(OK);
    }

    auto enqueue = [&](int it) -> int {
        hipLaunchKernelGGL(k_sc_beta, dim3(1), dim3(64), 0, s, w.st);
        hipLaunchKernelGGL(k_update_p, dim3(eg), dim3(BT), 0, s, w.st, w.p, w.r, w.v, n);
        MARIE_HIP(hipGetLastError());
        int st2 = OK;
        if (precond) st2 = apply(h, OP_P, 0, PREC_F64, w.p, ph, 1, s);
        if (!st2) st2 = apply(h, OP_AN, 0, PREC_F64, ph, w.v, 1, s);
        if (st2) return st2;
        hipLaunchKernelGGL(k_dot_rtv, dim3(G), dim3(BT), 0, s, w.st, w.rt, w.v, n, w.part, G);
        hipLaunchKernelGGL(k_sc_alpha, dim3(1), dim3(BT), 0, s, w.st, w.part, G);
        hipLaunchKernelGGL(k_update_s, dim3(G), dim3(BT), 0, s, w.st, w.s, w.r, w.v, x, ph, n, w.part, G);
        hipLaunchKernelGGL(k_sc_res, dim3(1), dim3(BT), 0, s, w.st, w.part, G, w.res);
        MARIE_HIP(hipGetLastError());
        if (precond) st2 = apply(h, OP_P, 0, PREC_F64, w.s, sh, 1, s);
        if (!st2) st2 = apply(h, OP_AN, 0, PREC_F64, sh, w.t, 1, s);
        if (st2) return st2;
        hipLaunchKernelGGL(k_dot_ts, dim3(G), dim3(BT), 0, s, w.st, w.t, w.s, n, w.part, G);
        hipLaunchKernelGGL(k_sc_omega, dim3(1), dim3(BT), 0, s, w.st, w.part, G);
        hipLaunchKernelGGL(k_update_r, dim3(G), dim3(BT), 0, s, w.st, x, sh, w.r, w.s, w.t, w.rt, n, w.part, G);
        hipLaunchKernelGGL(k_sc_r, dim3(1), dim3(BT), 0, s, w.st, w.part, G, w.res);
        MARIE_HIP(hipGetLastError());
        MARIE_HIP(hipMemcpyAsync(w.hst + (it & 1), w.st, sizeof(BiState), hipMemcpyDeviceToHost, s));
        MARIE_HIP(hipEventRecord(w.ev[it & 1], s));
        return OK;
    };
    int last = 0;
    if (!w.hst[0].done) {
        for (int it = 1; it <= maxit; ++it) {
            st = enqueue(it);
            if (st) return done(st);
            last = it;
            if (it >= 2) {
                MARIE_HIP(hipEventSynchronize(w.ev[(it - 1) & 1]));
                if (w.hst[(it - 1) & 1].done) break;
            }
        }
    }
    MARIE_HIP(hipMemcpyAsync(w.hst, w.st, sizeof(BiState), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipStreamSynchronize(s));
    const BiState fin = w.hst[0];
    (void)last;
    resvec.resize(fin.nres);
    MARIE_HIP(hipMemcpy(resvec.data(), w.res, fin.nres * sizeof(double), hipMemcpyDeviceToHost));
    *iter = fin.half / 2.0;
    // true residual
    st = apply(h, OP_AN, 0, PREC_F64, x, w.t, 1, s);
    if (st) return done(st);
    hipLaunchKernelGGL(k_sub, dim3(eg), dim3(BT), 0, s, w.t, b, w.t, n);
    hipLaunchKernelGGL(k_norm_partials, dim3(G), dim3(BT), 0, s, w.t, n, w.part, G);
    hipLaunchKernelGGL(k_sum1, dim3(1), dim3(BT), 0, s, w.part, G, w.scal + 1);
    MARIE_HIP(hipGetLastError());
    double tn = 0.0;
    MARIE_HIP(hipMemcpyAsync(&tn, w.scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipStreamSynchronize(s));
    *relres = std::sqrt(tn) / bnorm;
    *flag = (*relres <= tol) ? 0 : (fin.done == 2 ? 4 : 1);
    return done(OK);
}

}  // namespace marie

using namespace marie;

extern "C" int marie_vie_bicgstab_device(marie_vie_op* hh, const void* b, void* x, double tol, int maxit,
                                         int precond, int* flag, double* relres, double* iter, double* resvec,
                                         int resvec_cap, int* nres) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!b || !x || !flag || !relres || !iter || maxit < 1 || maxit > (1 << 24) || !(tol > 0) ||
        (precond != 0 && precond != 1)) {
        set_error("invalid bicgstab arguments (maxit >= 1, precond 0 or 1)");
        return MARIE_ERR_ARG;
    }
    if (!(h->precisions & PREC_F64)) {
        set_error("bicgstab needs an FP64 operator");
        return MARIE_ERR_STATE;
    }
    if (!h->have_fN || !h->have_mat || !h->have_idx) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) {
        *flag = 0; *relres = 0.0; *iter = 0.0;
        if (nres) *nres = 0;
        return MARIE_OK;
    }
    std::vector<double> rv;
    int st = bicgstab(h, (const c2<double>*)b, (c2<double>*)x, tol, maxit, precond, flag, relres, iter, rv);
    if (st) return st;
    if (resvec && resvec_cap > 0) {
        const int nc = (int)rv.size() < resvec_cap ? (int)rv.size() : resvec_cap;
        for (int i = 0; i < nc; ++i) resvec[i] = rv[i];
    }
    if (nres) *nres = (int)rv.size();
    return MARIE_OK;
}

extern "C" int marie_vie_bicgstab(marie_vie_op* hh, const double* b, double* x, double tol, int maxit, int precond,
                                  int* flag, double* relres, double* iter, double* resvec, int resvec_cap,
                                  int* nres) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!b || !x) { set_error("null vector"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)h->nidx;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(b[i]) || !std::isfinite(x[i])) {
            set_error("right-hand side or initial guess has non-finite entries");
            return MARIE_ERR_NONFINITE;
        }
    void *db = nullptr, *dxp = nullptr;
    if (n) {
        MARIE_HIP(hipMalloc(&db, n * sizeof(c2<double>)));
        if (hipMalloc(&dxp, n * sizeof(c2<double>)) != hipSuccess) {
            (void)hipFree(db);
            set_error("out of device memory");
            return MARIE_ERR_ALLOC;
        }
        MARIE_HIP(hipMemcpy(db, b, n * sizeof(c2<double>), hipMemcpyHostToDevice));
        MARIE_HIP(hipMemcpy(dxp, x, n * sizeof(c2<double>), hipMemcpyHostToDevice));
    }
    int st = marie_vie_bicgstab_device(hh, db, dxp, tol, maxit, precond, flag, relres, iter, resvec, resvec_cap,
                                       nres);
    if (st == MARIE_OK && n) {
        if (hipMemcpy(x, dxp, n * sizeof(c2<double>), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("D2H copy failed");
            st = MARIE_ERR_HIP;
        }
    }
    if (db) (void)hipFree(db);
    if (dxp) (void)hipFree(dxp);
    return st;
}
