// Device-resident restarted GMRES with MARIE's pgmres semantics, plus the
// mixed-precision variant (complex64 matvecs, FP64 Arnoldi, FP64 iterative
// refinement on the true residual).
//
// Semantics followed (src_numeric/src_iterative/pgmres.m:70-181,
// pgmres_iter.m:22-150): one-pass classical Gram-Schmidt h = V^H w,
// w -= V h; convergence test on the least-squares residual against
// tol*||b||; at the end of a cycle x += V y and the implicit residual
// r = V (rhs - H y); restart on resvec; flag/relres from the true residual.
// The least-squares problem is updated with Givens rotations (equal to the
// reference's from-scratch solve in exact arithmetic); the small dense
// algebra runs on the host in FP64, the O(n) vector work on the device with
// deterministic two-stage reductions (no atomics).
//
// The Arnoldi step needs nothing from the host: h, ||w|| and the next basis
// vector stay on the device, and each Hessenberg column is copied to pinned
// host memory behind an event. The host queues step k+1 before it waits for
// column k, so the device never idles on the host's Givens update; when
// column k meets the tolerance, the already queued step k+1 is discarded
// (its only outputs are basis column k+1 and Hessenberg column k+1, which
// the cycle's update never reads).
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"
#include "vie_reduce.hpp"
#include "vie_dense.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace marie {

using cd = std::complex<double>;
constexpr int KT = 256;      // threads per block
constexpr int KG = 1024;     // max blocks of a reduction
constexpr int KV = 8;        // basis vectors per pass of the multi-dot


// Vectors of the Krylov space are stored as c2<TV> (TV = double, or float for
// the complex64-basis mode); every product and sum accumulates in double.
template <typename TV> __device__ __forceinline__ double2 ldc(const c2<TV>* p) {
    const c2<TV> v = *p;
    return double2{(double)v.x, (double)v.y};
}
// a basis entry, read non-temporally: the basis is streamed once per
// multi-dot and once per update, and kept out of the caches it leaves the
// product's operator and work buffers there (FP64 GMRES(50) at 128^3:
// 155.9 -> 149.5 ms)
template <typename TV> __device__ __forceinline__ double2 ldb(const c2<TV>* p) {
    typedef TV v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return double2{(double)v.x, (double)v.y};
}

// partial[i * G + blk] = block sum of conj(V_i) * w over its range, i < k.
// Grid (G, ceil(k / KV)): blockIdx.y picks a group of KV basis vectors, so the
// groups are swept in parallel rather than one after another by each block.
template <typename TV>
__global__ __launch_bounds__(KT) void k_mdot(const c2<TV>* __restrict__ V, int64_t ldv, int k,
                                             const c2<TV>* __restrict__ w, int64_t n, c2<double>* __restrict__ partial,
                                             int G) {
    __shared__ double sh[KT / 64 * 2 * KV];
    const int i0 = blockIdx.y * KV;
    const int nv = (k - i0) < KV ? (k - i0) : KV;
    double acc[2 * KV];
#pragma unroll
    for (int q = 0; q < 2 * KV; ++q) acc[q] = 0.0;
    const int64_t st = (int64_t)G * KT;
    int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x;
    // two elements per thread per trip: 2 * (nv + 1) independent loads in flight
    for (; e + st < n; e += 2 * st) {
        const double2 w0 = ldc(w + e), w1 = ldc(w + e + st);
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            if (q < nv) {
                const c2<TV>* col = V + (i0 + q) * ldv + e;
                const double2 a = ldb(col), b = ldb(col + st);
                acc[2 * q] += a.x * w0.x + a.y * w0.y + b.x * w1.x + b.y * w1.y;
                acc[2 * q + 1] += a.x * w0.y - a.y * w0.x + b.x * w1.y - b.y * w1.x;
            }
        }
    }
    if (e < n) {
        const double2 wv = ldc(w + e);
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            if (q < nv) {
                const double2 a = ldb(V + (i0 + q) * ldv + e);
                acc[2 * q] += a.x * wv.x + a.y * wv.y;
                acc[2 * q + 1] += a.x * wv.y - a.y * wv.x;
            }
        }
    }
    block_reduce<2 * KV>(acc, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < nv; ++q)
            partial[(int64_t)(i0 + q) * G + blockIdx.x] = c2<double>{acc[2 * q], acc[2 * q + 1]};
}

// k_mdot for a complex64 basis with an even column stride: each thread reads
// two adjacent elements of w and of every basis vector with one 16-byte load
// (the last element of an odd length is added by thread 0 of block 0)
__global__ __launch_bounds__(KT) void k_mdot_f2(const c2<float>* __restrict__ V, int64_t ldv, int k,
                                                const c2<float>* __restrict__ w, int64_t n,
                                                c2<double>* __restrict__ partial, int G) {
    __shared__ double sh[KT / 64 * 2 * KV];
    const int64_t n2 = n >> 1;
    const float4* w4 = reinterpret_cast<const float4*>(w);
    const int i0 = blockIdx.y * KV;
    const int nv = (k - i0) < KV ? (k - i0) : KV;
    double acc[2 * KV];
#pragma unroll
    for (int q = 0; q < 2 * KV; ++q) acc[q] = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x; e < n2; e += (int64_t)G * KT) {
        const float4 wv = w4[e];
#pragma unroll
        for (int q = 0; q < KV; ++q) {
            if (q < nv) {
                const float4 a = reinterpret_cast<const float4*>(V + (i0 + q) * ldv)[e];
                acc[2 * q] += (double)a.x * wv.x + (double)a.y * wv.y + (double)a.z * wv.z + (double)a.w * wv.w;
                acc[2 * q + 1] += (double)a.x * wv.y - (double)a.y * wv.x + (double)a.z * wv.w - (double)a.w * wv.z;
            }
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const c2<float> wv = w[n - 1];
        for (int q = 0; q < nv; ++q) {
            const c2<float> a = V[(i0 + q) * ldv + n - 1];
            acc[2 * q] += (double)a.x * wv.x + (double)a.y * wv.y;
            acc[2 * q + 1] += (double)a.x * wv.y - (double)a.y * wv.x;
        }
    }
    block_reduce<2 * KV>(acc, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < nv; ++q)
            partial[(int64_t)(i0 + q) * G + blockIdx.x] = c2<double>{acc[2 * q], acc[2 * q + 1]};
}

// out[i] = sum_b partial[i*G + b], fixed order; one block per i
__global__ __launch_bounds__(KT) void k_reduce(const c2<double>* __restrict__ partial, int G,
                                               c2<double>* __restrict__ out) {
    __shared__ double sh[KT / 64 * 2];
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < G; b += KT) {
        const c2<double> p = partial[(int64_t)blockIdx.x * G + b];
        acc[0] += p.x;
        acc[1] += p.y;
    }
    block_reduce<2>(acc, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = c2<double>{acc[0], acc[1]};
}

// most basis columns one combination reads (restart + 1 of GMRES(restart))
constexpr int kMaxCols = 512;

// y = beta*y + csign * sum_{i<k} V_i c_i (k <= kMaxCols); optional partial ||y||^2 into npart[blk]
template <typename TY, typename TV>
__global__ __launch_bounds__(KT) void k_vcomb(c2<TY>* __restrict__ y, double beta, const c2<TV>* __restrict__ V,
                                              int64_t ldv, int k, const c2<double>* __restrict__ coef, double csign,
                                              int64_t n, double* __restrict__ npart, int G) {
    __shared__ c2<double> cs[kMaxCols];
    __shared__ double sh[KT / 64];
    for (int i = threadIdx.x; i < k; i += KT) cs[i] = coef[i];
    __syncthreads();
    double nrm[1] = {0.0};
    for (int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x; e < n; e += (int64_t)G * KT) {
        double2 acc = {0.0, 0.0};
        if (beta != 0.0) {
            const double2 yv = ldc(y + e);
            acc.x = beta * yv.x;
            acc.y = beta * yv.y;
        }
        for (int i = 0; i < k; ++i) {
            const double2 a = ldb(V + i * ldv + e);
            const c2<double> c = cs[i];
            acc.x += csign * (a.x * c.x - a.y * c.y);
            acc.y += csign * (a.x * c.y + a.y * c.x);
        }
        const c2<TY> out = {(TY)acc.x, (TY)acc.y};
        y[e] = out;
        const double ox = (double)out.x, oy = (double)out.y;
        nrm[0] += ox * ox + oy * oy;
    }
    if (npart) {
        block_reduce<1>(nrm, sh);
        if (threadIdx.x == 0) npart[blockIdx.x] = nrm[0];
    }
}

template <typename TV>
__global__ __launch_bounds__(KT) void k_norm2(const c2<TV>* __restrict__ v, int64_t n, double* __restrict__ npart,
                                              int G) {
    __shared__ double sh[KT / 64];
    double acc[1] = {0.0};
    for (int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x; e < n; e += (int64_t)G * KT) {
        const double2 a = ldc(v + e);
        acc[0] += a.x * a.x + a.y * a.y;
    }
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) npart[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(KT) void k_reduce_real(const double* __restrict__ p, int G, double* __restrict__ out) {
    __shared__ double sh[KT / 64];
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < G; b += KT) acc[0] += p[b];
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) out[0] = acc[0];
}

// v = w / ||w|| with ||w||^2 = sum of the G partials p (zero when the norm is
// zero: breakdown). Every block adds the partials in the same fixed order,
// so all use the same norm; block 0 stores it in *out (H(k+1,k)). One launch
// for the norm and the scaling.
template <typename TV>
__global__ __launch_bounds__(KT) void k_scale_by_norm(c2<TV>* __restrict__ v, const c2<TV>* __restrict__ w,
                                                      const double* __restrict__ p, int G,
                                                      c2<double>* __restrict__ out, int64_t n) {
    __shared__ double sh[KT / 64];
    __shared__ double nrm;
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < G; b += KT) acc[0] += p[b];
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) {
        nrm = sqrt(acc[0]);
        if (blockIdx.x == 0) *out = c2<double>{nrm, 0.0};
    }
    __syncthreads();
    const double d = nrm;
    const double a = d > 0.0 ? 1.0 / d : 0.0;
    for (int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x; e < n; e += (int64_t)gridDim.x * KT) {
        const double2 x = ldc(w + e);
        v[e] = c2<TV>{(TV)(a * x.x), (TV)(a * x.y)};
    }
}

// z = a*x + b*y (real scalars), any storage types
template <typename TZ, typename TX, typename TY>
__global__ void k_axpby(c2<TZ>* __restrict__ z, double a, const c2<TX>* __restrict__ x, double b,
                        const c2<TY>* __restrict__ y, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const double2 xv = ldc(x + e), yv = ldc(y + e);
        z[e] = c2<TZ>{(TZ)(a * xv.x + b * yv.x), (TZ)(a * xv.y + b * yv.y)};
    }
}

// ------------------------------------------------------------------ host side
struct Ws {
    void* V = nullptr;           // (m+1) x n of c2<TV>
    void* w = nullptr;           // n of c2<TV>
    void* rT = nullptr;          // n of c2<TV> (inner residual)
    c2<double>* r = nullptr;     // n, FP64 true residual
    c2<double>* t = nullptr;     // n, FP64 scratch / correction
    c2<float>* xs = nullptr;     // n (complex64 staging for FP64-basis mixed mode)
    c2<float>* ys = nullptr;
    c2<double>* pt = nullptr;    // n, FP64 preconditioner intermediate
    c2<double>* part = nullptr;  // (m+1) x KG
    c2<double>* h = nullptr;     // (m+1) x (m+1): Hessenberg columns, then update coefficients
    double* npart = nullptr;     // KG
    double* nout = nullptr;      // 1
    c2<double>* hp = nullptr;    // pinned host mirror of h
    hipEvent_t ev[2] = {nullptr, nullptr};
    void free_all() {
        void* p[] = {V, w, rT, r, t, xs, ys, pt, part, h, npart, nout};
        for (void* q : p)
            if (q) (void)hipFree(q);
        if (hp) (void)hipHostFree(hp);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

struct Ctx {
    Op* h;
    hipStream_t s;
    int64_t n;
    int G;
    int transp = 0;              // 1: solve with A^H (mv_AN's 'transp' branch)
    int precond = 0;             // 0 none, 1 right (A P), 2 left (P A): pgmres R1 / L1
    int64_t ldv = 0;             // column stride of the basis V (n, or n rounded up to even in mode 2)
    Ws ws;
};

static int grid_for(int64_t n) {
    int64_t g = (n + KT * 4 - 1) / (KT * 4);
    if (g < 1) g = 1;
    if (g > KG) g = KG;
    return (int)g;
}

template <typename TV> static int norm(Ctx& c, const c2<TV>* v, double* out) {
    hipLaunchKernelGGL(k_norm2<TV>, dim3(c.G), dim3(KT), 0, c.s, v, c.n, c.ws.npart, c.G);
    hipLaunchKernelGGL(k_reduce_real, dim3(1), dim3(KT), 0, c.s, c.ws.npart, c.G, c.ws.nout);
    MARIE_HIP(hipGetLastError());
    double v2 = 0.0;
    MARIE_HIP(hipMemcpyAsync(&v2, c.ws.nout, sizeof(double), hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipStreamSynchronize(c.s));
    *out = std::sqrt(v2);
    return OK;
}

// y = A x with vectors stored as TV; `f32op` selects the complex64 operator
template <typename TV> static int matvec(Ctx& c, const c2<TV>* x, c2<TV>* y, bool f32op) {
    if constexpr (sizeof(TV) == 4) {
        return apply(c.h, OP_AN, c.transp, PREC_F32, x, y, 1, c.s);
    } else {
        if (!f32op) return apply(c.h, OP_AN, c.transp, PREC_F64, x, y, 1, c.s);
        const int g = grid_for(c.n);
        hipLaunchKernelGGL((k_axpby<float, double, double>), dim3(g), dim3(KT), 0, c.s, c.ws.xs, 1.0, x, 0.0, x, c.n);
        int st = apply(c.h, OP_AN, c.transp, PREC_F32, c.ws.xs, c.ws.ys, 1, c.s);
        if (st) return st;
        hipLaunchKernelGGL((k_axpby<double, float, float>), dim3(g), dim3(KT), 0, c.s, y, 1.0, c.ws.ys, 0.0, c.ws.ys,
                           c.n);
        MARIE_HIP(hipGetLastError());
        return OK;
    }
}

// The cycle's operator: A, or with the preconditioner P = mv_AN(., 1./Mr,
// 1./Mr - 1) applied as pgmres_iter.m:64-96 does, A P (right) or P A (left)
template <typename TV> static int cycle_matvec(Ctx& c, const c2<TV>* x, c2<TV>* y, bool f32op) {
    if constexpr (sizeof(TV) == 8) {
        if (c.precond == 1) {
            const int st = apply(c.h, OP_P, 0, PREC_F64, x, c.ws.pt, 1, c.s);
            return st ? st : apply(c.h, OP_AN, 0, PREC_F64, c.ws.pt, y, 1, c.s);
        }
        if (c.precond == 2) {
            const int st = apply(c.h, OP_AN, 0, PREC_F64, x, c.ws.pt, 1, c.s);
            return st ? st : apply(c.h, OP_P, 0, PREC_F64, c.ws.pt, y, 1, c.s);
        }
    }
    return matvec<TV>(c, x, y, f32op);
}

// One restart cycle of pgmres_iter on A z = r (r: current residual, stored as TV).
// Accumulates the update into the FP64 vector x and leaves the implicit
// residual V (rhs - H y) in r. Appends |LS residual| values to resv.
template <typename TV>
static int gmres_cycle(Ctx& c, c2<double>* x, c2<TV>* r, int m, double abstol, bool f32op, int* kout,
                       std::vector<double>& resv) {
    const int64_t n = c.n;
    c2<TV>* V = reinterpret_cast<c2<TV>*>(c.ws.V);
    c2<TV>* w = reinterpret_cast<c2<TV>*>(c.ws.w);
    double beta = 0.0;
    int st = norm<TV>(c, r, &beta);
    if (st) return st;
    if (beta == 0.0) {
        *kout = 0;
        return OK;
    }
    // V0 = r / beta
    hipLaunchKernelGGL((k_axpby<TV, TV, TV>), dim3(c.G), dim3(KT), 0, c.s, V, 1.0 / beta, r, 0.0, r, n);
    const int ld = m + 1;                                // Hessenberg column stride (device and pinned)
    std::vector<cd> H((size_t)ld * m, cd(0, 0));         // column-major, original Hessenberg
    std::vector<cd> R((size_t)ld * m, cd(0, 0));         // rotated
    std::vector<double> cs(m);
    std::vector<cd> sn(m);
    std::vector<cd> g(m + 1, cd(0, 0));
    g[0] = cd(beta, 0);
    std::vector<cd> hcol(m + 1);

    // Arnoldi step k (1-based): w = A V(:,k), Hessenberg column k-1 on the
    // device, V(:,k+1) = w / H(k+1,k), column copied to pinned memory + event
    auto enqueue = [&](int k) -> int {
        int st2 = cycle_matvec<TV>(c, V + (int64_t)(k - 1) * c.ldv, w, f32op);
        if (st2) return st2;
        c2<double>* hc = c.ws.h + (size_t)(k - 1) * ld;
        // h = V(:,1:k)^H w  (classical Gram-Schmidt, pgmres_iter.m:106)
        if constexpr (sizeof(TV) == 4) {
            if ((c.ldv & 1) == 0)
                hipLaunchKernelGGL(k_mdot_f2, dim3(c.G, (k + KV - 1) / KV), dim3(KT), 0, c.s, V, c.ldv, k, w, n, c.ws.part,
                                   c.G);
            else
                hipLaunchKernelGGL(k_mdot<TV>, dim3(c.G, (k + KV - 1) / KV), dim3(KT), 0, c.s, V, c.ldv, k, w, n, c.ws.part,
                                   c.G);
        } else {
            hipLaunchKernelGGL(k_mdot<TV>, dim3(c.G, (k + KV - 1) / KV), dim3(KT), 0, c.s, V, c.ldv, k, w, n, c.ws.part,
                                   c.G);
        }
        hipLaunchKernelGGL(k_reduce, dim3(k), dim3(KT), 0, c.s, c.ws.part, c.G, hc);
        // w = w - V h with ||w||^2 partials (pgmres_iter.m:107, :109)
        hipLaunchKernelGGL((k_vcomb<TV, TV>), dim3(c.G), dim3(KT), 0, c.s, w, 1.0, V, c.ldv, k, hc, -1.0, n,
                           c.ws.npart, c.G);
        // H(k+1,k) = ||w||, V(:,k+1) = w / H(k+1,k)  (pgmres_iter.m:109-110)
        hipLaunchKernelGGL(k_scale_by_norm<TV>, dim3(c.G), dim3(KT), 0, c.s, V + (int64_t)k * c.ldv, w, c.ws.npart,
                           c.G, hc + k, n);
        MARIE_HIP(hipGetLastError());
        MARIE_HIP(hipMemcpyAsync(c.ws.hp + (size_t)(k - 1) * ld, hc, (k + 1) * sizeof(c2<double>),
                                 hipMemcpyDeviceToHost, c.s));
        MARIE_HIP(hipEventRecord(c.ws.ev[k & 1], c.s));
        return OK;
    };
    // Givens update with column k-1 (Arnoldi step k); true when the cycle stops there
    auto process = [&](int k, bool* stop) -> int {
        MARIE_HIP(hipEventSynchronize(c.ws.ev[k & 1]));
        const c2<double>* hh = c.ws.hp + (size_t)(k - 1) * ld;
        for (int i = 0; i <= k; ++i) hcol[i] = cd(hh[i].x, hh[i].y);
        const double hk1 = hh[k].x;
        for (int i = 0; i <= k; ++i) H[(size_t)(k - 1) * ld + i] = hcol[i];
        // apply previous rotations, then a new one zeroing hcol[k]
        for (int i = 0; i < k - 1; ++i) {
            const cd a = hcol[i], b = hcol[i + 1];
            hcol[i] = cs[i] * a + sn[i] * b;
            hcol[i + 1] = -std::conj(sn[i]) * a + cs[i] * b;
        }
        {
            const cd a = hcol[k - 1], b = hcol[k];
            const double na = std::abs(a), nb = std::abs(b);
            if (nb == 0.0) {
                cs[k - 1] = 1.0;
                sn[k - 1] = cd(0, 0);
            } else if (na == 0.0) {
                cs[k - 1] = 0.0;
                sn[k - 1] = std::conj(b) / nb;
            } else {
                const double tt = std::hypot(na, nb);
                cs[k - 1] = na / tt;
                sn[k - 1] = (a / na) * std::conj(b) / tt;
            }
            hcol[k - 1] = cs[k - 1] * a + sn[k - 1] * b;
            hcol[k] = cd(0, 0);
            const cd gk = g[k - 1];
            g[k - 1] = cs[k - 1] * gk;
            g[k] = -std::conj(sn[k - 1]) * gk;
        }
        for (int i = 0; i < k; ++i) R[(size_t)(k - 1) * ld + i] = hcol[i];
        const double res = std::abs(g[k]);
        resv.push_back(res);
        *stop = (res < abstol || hk1 == 0.0);
        return OK;
    };
    int k = 0;
    bool stop = false;
    for (int kk = 1; kk <= m; ++kk) {
        int st2 = enqueue(kk);
        if (st2) return st2;
        if (kk >= 2) {
            st2 = process(kk - 1, &stop);
            if (st2) return st2;
            if (stop) { k = kk - 1; break; }
        }
    }
    if (!stop) {
        int st2 = process(m, &stop);
        if (st2) return st2;
        k = m;
    }
    // y = R^{-1} g (upper triangular k x k)
    std::vector<cd> y(k);
    for (int i = k - 1; i >= 0; --i) {
        cd sacc = g[i];
        for (int j = i + 1; j < k; ++j) sacc -= R[(size_t)j * ld + i] * y[j];
        y[i] = sacc / R[(size_t)i * ld + i];
    }
    // LS residual vector res = rhs - Hc y (pgmres_iter.m:119) -> r = V(:,1:k+1) res (:150)
    std::vector<cd> rv(k + 1, cd(0, 0));
    rv[0] = cd(beta, 0);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i <= j + 1; ++i) rv[i] -= H[(size_t)j * ld + i] * y[j];
    // coefficients through the two slots after the m Hessenberg columns
    MARIE_HIP(hipStreamSynchronize(c.s));
    c2<double>* cy = c.ws.hp + (size_t)m * ld;
    c2<double>* cr = c.ws.hp + (size_t)m * ld + ld;
    for (int i = 0; i < k; ++i) cy[i] = c2<double>{y[i].real(), y[i].imag()};
    for (int i = 0; i <= k; ++i) cr[i] = c2<double>{rv[i].real(), rv[i].imag()};
    c2<double>* dy = c.ws.h + (size_t)m * ld;
    c2<double>* dr = c.ws.h + (size_t)m * ld + ld;
    MARIE_HIP(hipMemcpyAsync(dy, cy, k * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    MARIE_HIP(hipMemcpyAsync(dr, cr, (k + 1) * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    if (c.precond == 1) {
        // right preconditioning: x += P (V y)  (pgmres_iter.m:131-145)
        hipLaunchKernelGGL((k_vcomb<double, TV>), dim3(c.G), dim3(KT), 0, c.s, c.ws.t, 0.0, V, c.ldv, k, dy, 1.0, n,
                           (double*)nullptr, c.G);
        int st2 = apply(c.h, OP_P, 0, PREC_F64, c.ws.t, c.ws.pt, 1, c.s);
        if (st2) return st2;
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, x, 1.0, x, 1.0, c.ws.pt,
                           n);
    } else {
        hipLaunchKernelGGL((k_vcomb<double, TV>), dim3(c.G), dim3(KT), 0, c.s, x, 1.0, V, c.ldv, k, dy, 1.0, n,
                           (double*)nullptr, c.G);
    }
    hipLaunchKernelGGL((k_vcomb<TV, TV>), dim3(c.G), dim3(KT), 0, c.s, r, 0.0, V, c.ldv, k + 1, dr, 1.0, n,
                       (double*)nullptr, c.G);
    MARIE_HIP(hipGetLastError());
    MARIE_HIP(hipStreamSynchronize(c.s));
    *kout = k;
    return OK;
}

static int alloc_ws(Ctx& c, int m, int mode) {
    const int64_t n = c.n;
    const size_t ev = mode == 2 ? sizeof(c2<float>) : sizeof(c2<double>);
    auto A = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the Krylov workspace");
            return ERR_ALLOC;
        }
        return OK;
    };
    // complex64 basis columns start on 16-byte boundaries (even stride), so the
    // multi-dot reads two elements per 16-byte load
    c.ldv = mode == 2 ? ((n + 1) & ~(int64_t)1) : n;
    // mode 3 keeps its complex64 basis in the s-step workspace (vie_sstep.hip)
    int st = mode == 3 ? OK : A(&c.ws.V, (size_t)(m + 1) * c.ldv * ev);
    if (!st && mode != 3) st = A(&c.ws.w, n * ev);
    if (!st && (mode == 1 || mode == 2)) st = A(&c.ws.rT, n * ev);
    if (!st) st = A((void**)&c.ws.r, n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.t, n * sizeof(c2<double>));
    if (!st && mode == 1) st = A((void**)&c.ws.xs, n * sizeof(c2<float>));
    if (!st && mode == 1) st = A((void**)&c.ws.ys, n * sizeof(c2<float>));
    if (!st && c.precond) st = A((void**)&c.ws.pt, n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.part, (size_t)(m + 1) * KG * sizeof(c2<double>));
    // Hessenberg columns (m x (m+1)) + the two coefficient vectors of the update
    const size_t nh = (size_t)(m + 2) * (m + 1);
    if (!st) st = A((void**)&c.ws.h, nh * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.npart, KG * sizeof(double));
    if (!st) st = A((void**)&c.ws.nout, sizeof(double));
    if (!st && hipHostMalloc((void**)&c.ws.hp, nh * sizeof(c2<double>), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c.ws.hp = nullptr;
        set_error("pinned host allocation failed");
        st = ERR_ALLOC;
    }
    for (int e = 0; e < 2 && !st; ++e)
        if (hipEventCreateWithFlags(&c.ws.ev[e], hipEventDisableTiming) != hipSuccess) {
            set_error("hipEventCreate failed");
            st = ERR_HIP;
        }
    return st;
}

// Iterative refinement with one inner GMRES cycle per step: the cycle runs the
// complex64 operator with its basis stored as TV on the current FP64 residual,
// then x is updated and the true residual b - A x recomputed in FP64. So the
// residual is replaced by the FP64 one at every restart, and the complex64
// operator's error enters relative to the current residual, not to ||b||.
template <typename TV>
static int refine(Ctx& c, const c2<double>* b, c2<double>* x, int m, double tol, int maxit, double inner_tol,
                  double bnorm, std::vector<double>& resvec, int* p_out, int* outit_out) {
    c2<TV>* rT = reinterpret_cast<c2<TV>*>(c.ws.rT);
    int outit = 0, p = 0, st = OK;
    double rcur = resvec.back() * bnorm;
    while (outit < maxit && rcur > tol * bnorm) {
        ++outit;
        // the cycle stops at the global target (with a 10 % margin for the
        // complex64 operator's error), at most inner_tol relative, and reduces
        // the residual by at least a factor 2
        double rel = 0.9 * tol * bnorm / rcur;
        if (inner_tol > rel) rel = inner_tol;
        if (rel > 0.5) rel = 0.5;
        MARIE_HIP(hipMemsetAsync(c.ws.t, 0, c.n * sizeof(c2<double>), c.s));
        hipLaunchKernelGGL((k_axpby<TV, double, double>), dim3(c.G), dim3(KT), 0, c.s, rT, 1.0, c.ws.r, 0.0, c.ws.r,
                           c.n);
        std::vector<double> rv;
        st = gmres_cycle<TV>(c, c.ws.t, rT, m, rel * rcur, true, &p, rv);
        if (st) return st;
        // x += d ; r = b - A64 x
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, x, 1.0, x, 1.0, c.ws.t,
                           c.n);
        st = matvec<double>(c, x, c.ws.t, false);
        if (st) return st;
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, c.ws.r, 1.0, b, -1.0,
                           c.ws.t, c.n);
        MARIE_HIP(hipGetLastError());
        st = norm<double>(c, c.ws.r, &rcur);
        if (st) return st;
        // resvec: the inner LS residuals scaled to ||b||, then the refined true residual
        for (double v : rv) resvec.push_back(v / bnorm);
        resvec.push_back(rcur / bnorm);
    }
    *p_out = p;
    *outit_out = outit;
    return OK;
}

int gmres(Op* h, const c2<double>* b, c2<double>* x, int m, double tol, int maxit, int mode, double inner_tol,
          int* flag, double* relres, int* iter, std::vector<double>& resvec, int transp, int precond) {
    Ctx c;
    c.h = h;
    c.transp = transp;
    c.precond = precond;
    c.s = h->stream;
    c.n = h->nidx;
    c.G = grid_for(c.n);
    int st = alloc_ws(c, m, mode);
    if (st) { c.ws.free_all(); return st; }
    auto done = [&](int code) { (void)hipStreamSynchronize(c.s); c.ws.free_all(); return code; };
    double bnorm = 0.0;
    st = norm<double>(c, b, &bnorm);
    if (st) return done(st);
    if (bnorm == 0.0) {
        // zero right-hand side: x = 0 (jvie2solver.m:30-33 fast return)
        MARIE_HIP(hipMemsetAsync(x, 0, c.n * sizeof(c2<double>), c.s));
        *flag = 0; *relres = 0.0; iter[0] = 0; iter[1] = 0;
        resvec.assign(1, 0.0);
        return done(OK);
    }
    // r = b - A x0 (x0 may be zero)
    double xn = 0.0;
    st = norm<double>(c, x, &xn);
    if (st) return done(st);
    if (xn > 0.0) {
        st = matvec<double>(c, x, c.ws.t, false);
        if (st) return done(st);
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, c.ws.r, 1.0, b, -1.0,
                           c.ws.t, c.n);
    } else {
        MARIE_HIP(hipMemcpyAsync(c.ws.r, b, c.n * sizeof(c2<double>), hipMemcpyDeviceToDevice, c.s));
    }
    // left preconditioning works on P r and measures against ||P b||
    // (pgmres.m:122-142); flag and relres stay on the true residual
    double bnorm_sys = bnorm;
    if (c.precond == 2) {
        st = apply(c.h, OP_P, 0, PREC_F64, c.ws.r, c.ws.pt, 1, c.s);
        if (st) return done(st);
        MARIE_HIP(hipMemcpyAsync(c.ws.r, c.ws.pt, c.n * sizeof(c2<double>), hipMemcpyDeviceToDevice, c.s));
        st = apply(c.h, OP_P, 0, PREC_F64, b, c.ws.pt, 1, c.s);
        if (st) return done(st);
        st = norm<double>(c, c.ws.pt, &bnorm_sys);
        if (st) return done(st);
        if (bnorm_sys == 0.0) {
            set_error("the preconditioner maps the right-hand side to zero");
            return done(ERR_ARG);
        }
    }
    double rn = 0.0;
    st = norm<double>(c, c.ws.r, &rn);
    if (st) return done(st);
    resvec.assign(1, rn / bnorm_sys);
    int outit = 0, p = 0;
    if (mode == 0) {
        // pgmres.m:153-167
        while (resvec.back() > tol && outit < maxit) {
            ++outit;
            std::vector<double> rv;
            st = gmres_cycle<double>(c, x, c.ws.r, m, tol * bnorm_sys, false, &p, rv);
            if (st) return done(st);
            for (double v : rv) resvec.push_back(v / bnorm_sys);
            if (rv.empty()) break;
        }
    } else if (mode == 1) {
        st = refine<double>(c, b, x, m, tol, maxit, inner_tol, bnorm, resvec, &p, &outit);
        if (st) return done(st);
    } else if (mode == 2) {
        st = refine<float>(c, b, x, m, tol, maxit, inner_tol, bnorm, resvec, &p, &outit);
        if (st) return done(st);
    } else {
        st = sstep_refine(c.h, c.s, c.transp, b, x, c.ws.r, c.ws.t, bnorm, rn, m, tol, maxit, inner_tol, resvec, &p,
                          &outit);
        if (st) return done(st);
    }
    iter[0] = p;
    iter[1] = outit;
    if (mode != 0) {
        // the refinement ends on the FP64 true residual of the returned x
        // (r = b - A x, its norm the last resvec entry): pgmres.m:175's relres
        *relres = resvec.back();
        *flag = (*relres <= tol) ? 0 : 1;
        return done(OK);
    }
    // true residual (pgmres.m:175)
    st = matvec<double>(c, x, c.ws.t, false);
    if (st) return done(st);
    hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, c.ws.t, 1.0, b, -1.0, c.ws.t,
                       c.n);
    double tn = 0.0;
    st = norm<double>(c, c.ws.t, &tn);
    if (st) return done(st);
    *relres = tn / bnorm;
    *flag = (*relres <= tol) ? 0 : 1;
    return done(OK);
}

// ------------------------------------------------------------ GMRES with deflated restarts
// pgmresDR.m:73-282 / pgmresDR_iter.m:23-142 (GCRO-DR): between cycles a pair
// U, C = A U (C orthonormal, k columns) spanned by the harmonic Ritz vectors
// of smallest magnitude is kept; each cycle's Arnoldi vectors are made
// orthogonal to C as well; the update minimises the residual over
// span([U V_p]) through H2 = [D B; 0 H].
//
// Device layout: one array W of kmax + m + 1 columns holds C in its last k of
// the first kmax slots followed by V, so [C V] is one contiguous block for the
// multi-dot and the combinations. Since V is built orthogonal to C, the
// reference's two projections (B = C^H w, then H = V^H (w - C B)) are taken
// in one classical Gram-Schmidt pass over [C V] (equal in exact arithmetic).
// The small dense algebra (least squares, thin QR, the harmonic Ritz
// eigenproblems) runs on the host in FP64 (vie_dense.hpp).

// Out[:, j0 + q] (+)= sum_i In[:, i] coef[(j0 + q) * nin + i], q < KO, nin <= 128
template <int KO>
__global__ __launch_bounds__(KT) void k_mcomb(const c2<double>* __restrict__ In, int nin,
                                              const c2<double>* __restrict__ coef, int nout, int j0,
                                              c2<double>* __restrict__ Out, int accumulate, int64_t n) {
    __shared__ c2<double> cs[KO * 128];
    const int nq = (nout - j0) < KO ? (nout - j0) : KO;
    for (int t = threadIdx.x; t < nq * nin; t += KT) cs[t] = coef[(int64_t)j0 * nin + t];
    __syncthreads();
    for (int64_t e = blockIdx.x * (int64_t)KT + threadIdx.x; e < n; e += (int64_t)gridDim.x * KT) {
        double ar[KO], ai[KO];
#pragma unroll
        for (int q = 0; q < KO; ++q) ar[q] = ai[q] = 0.0;
        for (int i = 0; i < nin; ++i) {
            const c2<double> v = In[(int64_t)i * n + e];
#pragma unroll
            for (int q = 0; q < KO; ++q) {
                if (q < nq) {
                    const c2<double> c = cs[q * nin + i];
                    ar[q] += v.x * c.x - v.y * c.y;
                    ai[q] += v.x * c.y + v.y * c.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KO; ++q) {
            if (q < nq) {
                c2<double>* o = Out + (int64_t)(j0 + q) * n + e;
                if (accumulate) {
                    const c2<double> old = *o;
                    *o = c2<double>{old.x + ar[q], old.y + ai[q]};
                } else {
                    *o = c2<double>{ar[q], ai[q]};
                }
            }
        }
    }
}

namespace {

constexpr int DR_MAXK = 32;

struct DrWs {
    c2<double>*U = nullptr, *T1 = nullptr, *T2 = nullptr, *xacc = nullptr, *coef = nullptr;
    void free_all() {
        void* p[] = {U, T1, T2, xacc, coef};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
};

using dense::cd;
using dense::Mat;

// coefficients (host, column-major nin x nout) -> device, then Out = In coef
static int mcomb(Ctx& c, DrWs& d, const c2<double>* In, int nin, const Mat& coef, int nout, c2<double>* Out,
                 bool accumulate) {
    if (nout == 0 || nin == 0) return OK;
    MARIE_HIP(hipStreamSynchronize(c.s));      // the pinned staging area is reused
    c2<double>* hc = c.ws.hp;
    for (size_t i = 0; i < (size_t)nin * nout; ++i) hc[i] = c2<double>{coef[i].real(), coef[i].imag()};
    MARIE_HIP(hipMemcpyAsync(d.coef, hc, (size_t)nin * nout * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    for (int j0 = 0; j0 < nout; j0 += 8)
        hipLaunchKernelGGL(k_mcomb<8>, dim3(c.G), dim3(KT), 0, c.s, In, nin, d.coef, nout, j0, Out,
                           accumulate ? 1 : 0, c.n);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// g = [cols]^H v (ncols entries) to the host
static int multidot(Ctx& c, const c2<double>* cols, int ncols, const c2<double>* v, std::vector<cd>& g) {
    hipLaunchKernelGGL(k_mdot<double>, dim3(c.G, (ncols + KV - 1) / KV), dim3(KT), 0, c.s, cols, c.n, ncols, v, c.n,
                       c.ws.part, c.G);
    hipLaunchKernelGGL(k_reduce, dim3(ncols), dim3(KT), 0, c.s, c.ws.part, c.G, c.ws.h);
    MARIE_HIP(hipGetLastError());
    MARIE_HIP(hipMemcpyAsync(c.ws.hp, c.ws.h, ncols * sizeof(c2<double>), hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipStreamSynchronize(c.s));
    g.resize(ncols);
    for (int i = 0; i < ncols; ++i) g[i] = cd(c.ws.hp[i].x, c.ws.hp[i].y);
    return OK;
}

// y += sign * cols coef (coef on the host)
static int vcomb_host(Ctx& c, c2<double>* y, const c2<double>* cols, int ncols, const std::vector<cd>& coef,
                      double sign) {
    if (ncols == 0) return OK;
    MARIE_HIP(hipStreamSynchronize(c.s));
    for (int i = 0; i < ncols; ++i) c.ws.hp[i] = c2<double>{coef[i].real(), coef[i].imag()};
    MARIE_HIP(hipMemcpyAsync(c.ws.h, c.ws.hp, ncols * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL((k_vcomb<double, double>), dim3(c.G), dim3(KT), 0, c.s, y, 1.0, cols, c.n, ncols, c.ws.h,
                       sign, c.n, (double*)nullptr, c.G);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// One pgmresDR_iter cycle on r with the recycled block Cb (k columns right
// before V in memory). Returns p, H ((p+1) x p), B (k x p); appends the LS
// residuals to resv.
static int dr_cycle(Ctx& c, c2<double>* Cb, int k, c2<double>* V, const c2<double>* r, int m, double abstol,
                    int* pout, Mat& H, Mat& B, std::vector<double>& resv) {
    const int64_t n = c.n;
    c2<double>* w = reinterpret_cast<c2<double>*>(c.ws.w);
    double beta = 0.0;
    int st = norm<double>(c, r, &beta);
    if (st) return st;
    hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, V, 1.0 / beta, r, 0.0, r, n);
    const int ld = m + 1;
    Mat Hf((size_t)ld * m, 0.0), Bf((size_t)(k > 0 ? k : 1) * m, 0.0);
    std::vector<double> cs(m);
    std::vector<cd> sn(m), g(m + 1, 0.0), col(k + m + 1);
    g[0] = beta;
    int p = m;
    for (int j = 1; j <= m; ++j) {
        st = cycle_matvec<double>(c, V + (int64_t)(j - 1) * n, w, false);
        if (st) return st;
        const int nc = k + j;                 // [C V_1..j]
        hipLaunchKernelGGL(k_mdot<double>, dim3(c.G, (nc + KV - 1) / KV), dim3(KT), 0, c.s, Cb, n, nc, w, n, c.ws.part,
                           c.G);
        hipLaunchKernelGGL(k_reduce, dim3(nc), dim3(KT), 0, c.s, c.ws.part, c.G, c.ws.h);
        hipLaunchKernelGGL((k_vcomb<double, double>), dim3(c.G), dim3(KT), 0, c.s, w, 1.0, Cb, n, nc, c.ws.h, -1.0,
                           n, c.ws.npart, c.G);
        hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(c.G), dim3(KT), 0, c.s, V + (int64_t)j * n, w, c.ws.npart,
                           c.G, c.ws.h + nc, n);
        MARIE_HIP(hipGetLastError());
        MARIE_HIP(hipMemcpyAsync(c.ws.hp, c.ws.h, (nc + 1) * sizeof(c2<double>), hipMemcpyDeviceToHost, c.s));
        MARIE_HIP(hipStreamSynchronize(c.s));
        for (int i = 0; i < k; ++i) Bf[(size_t)(j - 1) * k + i] = cd(c.ws.hp[i].x, c.ws.hp[i].y);
        for (int i = 0; i <= j; ++i) {
            col[i] = cd(c.ws.hp[k + i].x, c.ws.hp[k + i].y);
            Hf[(size_t)(j - 1) * ld + i] = col[i];
        }
        // Givens update of the least-squares problem H y = beta e1
        for (int i = 0; i < j - 1; ++i) {
            const cd a = col[i], b = col[i + 1];
            col[i] = cs[i] * a + sn[i] * b;
            col[i + 1] = -std::conj(sn[i]) * a + cs[i] * b;
        }
        double cc;
        cd ss;
        dense::givens(col[j - 1], col[j], cc, ss);
        cs[j - 1] = cc;
        sn[j - 1] = ss;
        const cd gk = g[j - 1];
        g[j - 1] = cc * gk;
        g[j] = -std::conj(ss) * gk;
        const double res = std::abs(g[j]);
        resv.push_back(res);
        if (res < abstol) {
            p = j;
            break;
        }
    }
    H.assign((size_t)(p + 1) * p, 0.0);
    for (int j = 0; j < p; ++j)
        for (int i = 0; i <= p; ++i) H[(size_t)j * (p + 1) + i] = Hf[(size_t)j * ld + i];
    B.assign((size_t)k * p, 0.0);
    for (size_t i = 0; i < (size_t)k * p; ++i) B[i] = Bf[i];
    *pout = p;
    return OK;
}

// indices of the eigenvalues sorted by magnitude (ascending, stable)
static std::vector<int> by_magnitude(const std::vector<cd>& v) {
    std::vector<int> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) o[i] = (int)i;
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return std::abs(v[a]) < std::abs(v[b]); });
    return o;
}

}  // namespace

int gmres_dr(Op* h, const c2<double>* b, c2<double>* x, int m, double tol, int maxit, int ritz, int precond,
             int* flag, double* relres, int* iter, std::vector<double>& resvec) {
    Ctx c;
    c.h = h;
    c.precond = precond;
    c.s = h->stream;
    c.n = h->nidx;
    c.G = grid_for(c.n);
    c.ldv = c.n;
    const int64_t n = c.n;
    // pgmresDR.m:112-116
    ritz = ritz < 1 ? std::min(10, (m + 1) / 2) : std::min(ritz, m - 1);
    const int kmax = ritz;
    DrWs d;
    auto A = [&](void** p, size_t bytes) -> int {
        if (bytes == 0) return OK;
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the GMRES-DR workspace");
            return ERR_ALLOC;
        }
        return OK;
    };
    const size_t vb = (size_t)n * sizeof(c2<double>);
    const int ncol = kmax + m + 1;
    int st = A(&c.ws.V, (size_t)ncol * vb);
    if (!st) st = A(&c.ws.w, vb);
    if (!st) st = A((void**)&c.ws.r, vb);
    if (!st) st = A((void**)&c.ws.t, vb);
    if (!st && precond) st = A((void**)&c.ws.pt, vb);
    if (!st) st = A((void**)&c.ws.part, (size_t)ncol * KG * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.h, (size_t)(ncol + 1) * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.npart, KG * sizeof(double));
    if (!st) st = A((void**)&c.ws.nout, sizeof(double));
    if (!st) st = A((void**)&d.U, (size_t)kmax * vb);
    if (!st) st = A((void**)&d.T1, (size_t)kmax * vb);
    if (!st) st = A((void**)&d.T2, (size_t)kmax * vb);
    if (!st) st = A((void**)&d.xacc, vb);
    const size_t ncoef = (size_t)ncol * (kmax > 0 ? kmax : 1) + ncol;
    if (!st) st = A((void**)&d.coef, ncoef * sizeof(c2<double>));
    if (!st && hipHostMalloc((void**)&c.ws.hp, ncoef * sizeof(c2<double>), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c.ws.hp = nullptr;
        set_error("pinned host allocation failed");
        st = ERR_ALLOC;
    }
    auto done = [&](int code) {
        (void)hipStreamSynchronize(c.s);
        c.ws.free_all();
        d.free_all();
        return code;
    };
    if (st) return done(st);
    c2<double>* W = reinterpret_cast<c2<double>*>(c.ws.V);
    c2<double>* V = W + (int64_t)kmax * n;
    c2<double>* r = c.ws.r;
    auto fail = [&](const char* msg) {
        set_error(msg);
        return done(ERR_STATE);
    };

    double bnorm_true = 0.0;
    st = norm<double>(c, b, &bnorm_true);
    if (st) return done(st);
    if (bnorm_true == 0.0) {
        MARIE_HIP(hipMemsetAsync(x, 0, vb, c.s));
        *flag = 0; *relres = 0.0; iter[0] = 0; iter[1] = 0;
        resvec.assign(1, 0.0);
        return done(OK);
    }
    // r = b - A x0 (pgmresDR.m:150); left preconditioning works on P r, P b
    double xn = 0.0;
    st = norm<double>(c, x, &xn);
    if (st) return done(st);
    if (xn > 0.0) {
        st = matvec<double>(c, x, c.ws.t, false);
        if (st) return done(st);
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, r, 1.0, b, -1.0, c.ws.t,
                           n);
    } else {
        MARIE_HIP(hipMemcpyAsync(r, b, vb, hipMemcpyDeviceToDevice, c.s));
    }
    double bnorm = bnorm_true;
    if (precond == 2) {
        st = apply(h, OP_P, 0, PREC_F64, r, c.ws.pt, 1, c.s);
        if (st) return done(st);
        MARIE_HIP(hipMemcpyAsync(r, c.ws.pt, vb, hipMemcpyDeviceToDevice, c.s));
        st = apply(h, OP_P, 0, PREC_F64, b, c.ws.pt, 1, c.s);
        if (st) return done(st);
        st = norm<double>(c, c.ws.pt, &bnorm);
        if (st) return done(st);
        if (bnorm == 0.0) return fail("the preconditioner maps the right-hand side to zero");
    }
    MARIE_HIP(hipMemsetAsync(d.xacc, 0, vb, c.s));
    double rn = 0.0;
    st = norm<double>(c, r, &rn);
    if (st) return done(st);
    resvec.assign(1, rn / bnorm);
    const double abstol = tol * bnorm;

    // first cycle: GMRES, y = H \ (V^H r), harmonic Ritz vectors of H (pgmresDR.m:165-201)
    int p = 0, k = 0, outit = 0;
    Mat H, B;
    std::vector<double> rv;
    st = dr_cycle(c, W + (int64_t)kmax * n, 0, V, r, m, abstol, &p, H, B, rv);
    if (st) return done(st);
    for (double v : rv) resvec.push_back(v / bnorm);
    {
        std::vector<cd> g, y;
        st = multidot(c, V, p + 1, r, g);
        if (st) return done(st);
        if (!dense::lstsq(H, p + 1, p, g, y)) return fail("GMRES-DR: singular least-squares problem");
        st = vcomb_host(c, d.xacc, V, p, y, 1.0);
        if (st) return done(st);
        std::vector<cd> z(p + 1, 0.0);
        for (int j = 0; j < p; ++j)
            for (int i = 0; i <= p; ++i) z[i] += H[(size_t)j * (p + 1) + i] * y[j];
        st = vcomb_host(c, r, V, p + 1, z, -1.0);
        if (st) return done(st);
    }
    if (resvec.back() > tol && p > 1) {
        k = ritz > p - 1 ? p - 1 : ritz;
        // getHarmVecs1.m:16-28: eig(H_p + h^2 H_p^{-H} e_p e_p^T), k smallest
        Mat Hp((size_t)p * p), Z((size_t)p, 0.0);
        for (int j = 0; j < p; ++j)
            for (int i = 0; i < p; ++i) Hp[(size_t)j * p + i] = H[(size_t)j * (p + 1) + i];
        Mat HpH((size_t)p * p);
        for (int j = 0; j < p; ++j)
            for (int i = 0; i < p; ++i) HpH[(size_t)j * p + i] = std::conj(Hp[(size_t)i * p + j]);
        // z = H_p^{-H} e_p
        std::vector<cd> ep(p, 0.0), zz;
        ep[p - 1] = 1.0;
        if (!dense::lstsq(HpH, p, p, ep, zz)) return fail("GMRES-DR: singular Hessenberg block");
        const cd hh = H[(size_t)(p - 1) * (p + 1) + p];
        Mat Mt = Hp;
        for (int i = 0; i < p; ++i) Mt[(size_t)(p - 1) * p + i] += hh * hh * zz[i];
        std::vector<cd> vals;
        Mat vecs;
        if (!dense::eig(Mt, p, vals, vecs)) return fail("GMRES-DR: eigenvalue iteration did not converge");
        const std::vector<int> o = by_magnitude(vals);
        Mat P((size_t)p * k);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < p; ++i) P[(size_t)j * p + i] = vecs[(size_t)o[j] * p + i];
        // [Q, R] = qr(H P, 0); C = V Q; U = V_p P R^{-1}
        Mat HP((size_t)(p + 1) * k, 0.0), Q, R;
        for (int j = 0; j < k; ++j)
            for (int l = 0; l < p; ++l)
                for (int i = 0; i <= p; ++i) HP[(size_t)j * (p + 1) + i] += H[(size_t)l * (p + 1) + i] * P[(size_t)j * p + l];
        dense::qr_thin(HP, p + 1, k, Q, R);
        if (!dense::right_solve_upper(P, p, k, R)) return fail("GMRES-DR: rank-deficient recycled block");
        st = mcomb(c, d, V, p, P, k, d.U, false);
        if (!st) st = mcomb(c, d, V, p + 1, Q, k, W + (int64_t)(kmax - k) * n, false);
        if (st) return done(st);
    }
    // later cycles (pgmresDR.m:205-251)
    while (resvec.back() > tol && outit < maxit) {
        ++outit;
        c2<double>* Cb = W + (int64_t)(kmax - k) * n;
        rv.clear();
        st = dr_cycle(c, Cb, k, V, r, m, abstol, &p, H, B, rv);
        if (st) return done(st);
        for (double v : rv) resvec.push_back(v / bnorm);
        // D = diag(1 ./ ||U_i||)
        std::vector<double> dinv(k);
        for (int i = 0; i < k; ++i) {
            double ni = 0.0;
            st = norm<double>(c, d.U + (int64_t)i * n, &ni);
            if (st) return done(st);
            if (!(ni > 0.0)) return fail("GMRES-DR: zero recycled vector");
            dinv[i] = 1.0 / ni;
        }
        // H2 = [D B; 0 H], (k + p + 1) x (k + p)
        const int r2 = k + p + 1, c2n = k + p;
        Mat H2((size_t)r2 * c2n, 0.0);
        for (int i = 0; i < k; ++i) H2[(size_t)i * r2 + i] = dinv[i];
        for (int j = 0; j < p; ++j) {
            for (int i = 0; i < k; ++i) H2[(size_t)(k + j) * r2 + i] = B[(size_t)j * k + i];
            for (int i = 0; i <= p; ++i) H2[(size_t)(k + j) * r2 + k + i] = H[(size_t)j * (p + 1) + i];
        }
        std::vector<cd> g, y;
        st = multidot(c, Cb, r2, r, g);
        if (st) return done(st);
        if (!dense::lstsq(H2, r2, c2n, g, y)) return fail("GMRES-DR: singular least-squares problem");
        // x += [U D, V_p] y ; r -= [C V] H2 y
        std::vector<cd> yu(k), yv(p), z(r2, 0.0);
        for (int i = 0; i < k; ++i) yu[i] = y[i] * dinv[i];
        for (int i = 0; i < p; ++i) yv[i] = y[k + i];
        st = vcomb_host(c, d.xacc, d.U, k, yu, 1.0);
        if (!st) st = vcomb_host(c, d.xacc, V, p, yv, 1.0);
        for (int j = 0; j < c2n; ++j)
            for (int i = 0; i < r2; ++i) z[i] += H2[(size_t)j * r2 + i] * y[j];
        if (!st) st = vcomb_host(c, r, Cb, r2, z, -1.0);
        if (st) return done(st);
        if (!(resvec.back() > tol) || k == 0) continue;
        // getHarmVecs2.m:18-35 with the normalised U = U_raw D
        const int m2 = k + p;
        Mat Mcv((size_t)r2 * k);                        // [C V]^H U (normalised)
        for (int i = 0; i < k; ++i) {
            st = multidot(c, Cb, r2, d.U + (int64_t)i * n, g);
            if (st) return done(st);
            for (int l = 0; l < r2; ++l) Mcv[(size_t)i * r2 + l] = g[l] * dinv[i];
        }
        Mat Am((size_t)(m2 + 1) * m2, 0.0);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < r2; ++i) Am[(size_t)j * (m2 + 1) + i] = Mcv[(size_t)j * r2 + i];
        for (int j = k; j < m2; ++j) Am[(size_t)j * (m2 + 1) + j] = 1.0;
        Mat AH((size_t)m2 * m2, 0.0), BH((size_t)m2 * m2, 0.0);
        for (int j = 0; j < m2; ++j)
            for (int i = 0; i < m2; ++i) {
                cd sa = 0.0, sb = 0.0;
                for (int l = 0; l < r2; ++l) {
                    const cd hc = std::conj(H2[(size_t)i * r2 + l]);
                    sa += hc * Am[(size_t)j * (m2 + 1) + l];
                    sb += hc * H2[(size_t)j * r2 + l];
                }
                AH[(size_t)j * m2 + i] = sa;
                BH[(size_t)j * m2 + i] = sb;
            }
        // eig(A, B) with B = H2^H H2 Hermitian positive definite: eig(B^{-1} A)
        if (!dense::chol_solve(BH, m2, AH, m2)) return fail("GMRES-DR: H2 lost full column rank");
        std::vector<cd> vals;
        Mat vecs;
        if (!dense::eig(AH, m2, vals, vecs)) return fail("GMRES-DR: eigenvalue iteration did not converge");
        const std::vector<int> o = by_magnitude(vals);
        Mat P((size_t)m2 * k);
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < m2; ++i) P[(size_t)j * m2 + i] = vecs[(size_t)o[m2 - k + j] * m2 + i];
        // [Q, R] = qr(H2 P, 0); C = [C V] Q; U = [U D, V_p] P R^{-1}
        Mat HP((size_t)r2 * k, 0.0), Q, R;
        for (int j = 0; j < k; ++j)
            for (int l = 0; l < m2; ++l)
                for (int i = 0; i < r2; ++i) HP[(size_t)j * r2 + i] += H2[(size_t)l * r2 + i] * P[(size_t)j * m2 + l];
        dense::qr_thin(HP, r2, k, Q, R);
        if (!dense::right_solve_upper(P, m2, k, R)) return fail("GMRES-DR: rank-deficient recycled block");
        Mat PU((size_t)k * k), PV((size_t)p * k);
        for (int j = 0; j < k; ++j) {
            for (int i = 0; i < k; ++i) PU[(size_t)j * k + i] = P[(size_t)j * m2 + i] * dinv[i];
            for (int i = 0; i < p; ++i) PV[(size_t)j * p + i] = P[(size_t)j * m2 + k + i];
        }
        st = mcomb(c, d, d.U, k, PU, k, d.T1, false);
        if (!st) st = mcomb(c, d, V, p, PV, k, d.T1, true);
        if (!st) st = mcomb(c, d, Cb, r2, Q, k, d.T2, false);
        if (st) return done(st);
        MARIE_HIP(hipMemcpyAsync(Cb, d.T2, (size_t)k * vb, hipMemcpyDeviceToDevice, c.s));
        std::swap(d.U, d.T1);
    }
    // x = x0 + R1(x) (pgmresDR.m:254-269)
    if (precond == 1) {
        st = apply(h, OP_P, 0, PREC_F64, d.xacc, c.ws.pt, 1, c.s);
        if (st) return done(st);
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, x, 1.0, x, 1.0, c.ws.pt,
                           n);
    } else {
        hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, x, 1.0, x, 1.0, d.xacc,
                           n);
    }
    MARIE_HIP(hipGetLastError());
    iter[0] = p;
    iter[1] = outit;
    st = matvec<double>(c, x, c.ws.t, false);
    if (st) return done(st);
    hipLaunchKernelGGL((k_axpby<double, double, double>), dim3(c.G), dim3(KT), 0, c.s, c.ws.t, 1.0, b, -1.0, c.ws.t,
                       n);
    double tn = 0.0;
    st = norm<double>(c, c.ws.t, &tn);
    if (st) return done(st);
    *relres = tn / bnorm_true;
    *flag = (*relres <= tol) ? 0 : 1;
    return done(OK);
}

}  // namespace marie

using namespace marie;

static int gmres_entry(marie_vie_op* hh, const void* b, void* x, int restart, double tol, int maxit, int mode,
                       double inner_tol, int* flag, double* relres, int* iter, double* resvec, int resvec_cap, int* nres,
                       int transp, int precond = 0) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!b || !x || !flag || !relres || !iter || restart < 1 || restart > kMaxCols - 1 || maxit < 1 || !(tol > 0) ||
        (mode < 0 || mode > 3) || (mode == 3 && restart > 503)) {
        set_error("invalid gmres arguments (restart must be 1..511, 1..503 in mode 3; mode 0..3)");
        return MARIE_ERR_ARG;
    }
    if (!(h->precisions & PREC_F64) || (mode != 0 && !(h->precisions & PREC_F32))) {
        set_error("gmres needs an FP64 operator (and FP32 for modes 1-3)");
        return MARIE_ERR_STATE;
    }
    if (!h->have_fN || !h->have_mat || !h->have_idx) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) {
        *flag = 0; *relres = 0.0; iter[0] = iter[1] = 0;
        if (nres) *nres = 0;
        return MARIE_OK;
    }
    if (precond != 0 && (mode != 0 || transp)) {
        set_error("preconditioned GMRES runs in mode 0 on the forward system");
        return MARIE_ERR_ARG;
    }
    std::vector<double> rv;
    int st = gmres(h, (const c2<double>*)b, (c2<double>*)x, restart, tol, maxit, mode, inner_tol, flag, relres,
                   iter, rv, transp, precond);
    if (st) return st;
    if (resvec && resvec_cap > 0) {
        const int nc = (int)rv.size() < resvec_cap ? (int)rv.size() : resvec_cap;
        for (int i = 0; i < nc; ++i) resvec[i] = rv[i];
    }
    if (nres) *nres = (int)rv.size();
    return MARIE_OK;
}

extern "C" int marie_vie_gmres_device(marie_vie_op* hh, const void* b, void* x, int restart, double tol, int maxit,
                                      int mode, double inner_tol, int* flag, double* relres, int* iter,
                                      double* resvec, int resvec_cap, int* nres) {
    return gmres_entry(hh, b, x, restart, tol, maxit, mode, inner_tol, flag, relres, iter, resvec, resvec_cap, nres,
                       0);
}

extern "C" int marie_vie_gmres_adjoint_device(marie_vie_op* hh, const void* g, void* y, int restart, double tol,
                                              int maxit, int mode, double inner_tol, int* flag, double* relres,
                                              int* iter, double* resvec, int resvec_cap, int* nres) {
    return gmres_entry(hh, g, y, restart, tol, maxit, mode, inner_tol, flag, relres, iter, resvec, resvec_cap, nres,
                       1);
}

extern "C" int marie_vie_gmres_precond_device(marie_vie_op* hh, const void* b, void* x, int restart, double tol,
                                              int maxit, int side, int* flag, double* relres, int* iter,
                                              double* resvec, int resvec_cap, int* nres) {
    if (side != MARIE_PRECOND_NONE && side != MARIE_PRECOND_RIGHT && side != MARIE_PRECOND_LEFT) {
        set_error("side must be MARIE_PRECOND_NONE, _RIGHT or _LEFT");
        return MARIE_ERR_ARG;
    }
    return gmres_entry(hh, b, x, restart, tol, maxit, 0, 0.0, flag, relres, iter, resvec, resvec_cap, nres, 0,
                       side);
}

// Host-buffer form of a device solve: upload b and x0, run, download x.
template <typename F>
static int host_solve(marie_vie_op* hh, const double* b, double* x, F&& solve) {
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
    int st = solve(db, dxp);
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

extern "C" int marie_vie_gmres(marie_vie_op* hh, const double* b, double* x, int restart, double tol, int maxit,
                               int mode, double inner_tol, int* flag, double* relres, int* iter, double* resvec,
                               int resvec_cap, int* nres) {
    return host_solve(hh, b, x, [&](void* db, void* dx) {
        return marie_vie_gmres_device(hh, db, dx, restart, tol, maxit, mode, inner_tol, flag, relres, iter, resvec,
                                      resvec_cap, nres);
    });
}

extern "C" int marie_vie_gmres_precond(marie_vie_op* hh, const double* b, double* x, int restart, double tol,
                                       int maxit, int side, int* flag, double* relres, int* iter, double* resvec,
                                       int resvec_cap, int* nres) {
    return host_solve(hh, b, x, [&](void* db, void* dx) {
        return marie_vie_gmres_precond_device(hh, db, dx, restart, tol, maxit, side, flag, relres, iter, resvec,
                                              resvec_cap, nres);
    });
}

extern "C" int marie_vie_gmres_dr_device(marie_vie_op* hh, const void* b, void* x, int restart, double tol,
                                         int maxit, int ritz, int side, int* flag, double* relres, int* iter,
                                         double* resvec, int resvec_cap, int* nres) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!b || !x || !flag || !relres || !iter || restart < 2 || restart > 64 || maxit < 1 || !(tol > 0) ||
        ritz > DR_MAXK || (side != MARIE_PRECOND_NONE && side != MARIE_PRECOND_RIGHT && side != MARIE_PRECOND_LEFT)) {
        set_error("invalid gmres_dr arguments (restart 2..64, ritz <= 32, side NONE/RIGHT/LEFT)");
        return MARIE_ERR_ARG;
    }
    if (!(h->precisions & PREC_F64)) {
        set_error("gmres_dr needs an FP64 operator");
        return MARIE_ERR_STATE;
    }
    if (!h->have_fN || !h->have_mat || !h->have_idx) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) {
        *flag = 0; *relres = 0.0; iter[0] = iter[1] = 0;
        if (nres) *nres = 0;
        return MARIE_OK;
    }
    std::vector<double> rv;
    int st = gmres_dr(h, (const c2<double>*)b, (c2<double>*)x, restart, tol, maxit, ritz, side, flag, relres, iter,
                      rv);
    if (st) return st;
    if (resvec && resvec_cap > 0) {
        const int nc = (int)rv.size() < resvec_cap ? (int)rv.size() : resvec_cap;
        for (int i = 0; i < nc; ++i) resvec[i] = rv[i];
    }
    if (nres) *nres = (int)rv.size();
    return MARIE_OK;
}

extern "C" int marie_vie_gmres_dr(marie_vie_op* hh, const double* b, double* x, int restart, double tol, int maxit,
                                  int ritz, int side, int* flag, double* relres, int* iter, double* resvec,
                                  int resvec_cap, int* nres) {
    return host_solve(hh, b, x, [&](void* db, void* dx) {
        return marie_vie_gmres_dr_device(hh, db, dx, restart, tol, maxit, ritz, side, flag, relres, iter, resvec,
                                         resvec_cap, nres);
    });
}
