// Restarted block GMRES for s right-hand sides of one operator (coil ports,
// excitations): one Krylov space of block vectors, one s-column product per
// block step (the fN reads of the z pass are shared by the s columns).
//
// The reference solves ports one at a time (VSIE_Solver.m:387-407,
// fVIEsolver_Assembly.m:185-197 build one pgmres per column); this is the
// batched counterpart, with per-column flag/relres semantics of pgmres.m:175.
//
// Per cycle, on R = B - A X (FP64):
//   R = V_0 S_0                       shifted CholeskyQR3 (Gram on device, s x s Cholesky on host)
//   block step j: W = A V_j           multi-RHS product (complex128, or complex64 in mode 1)
//                 H_{0:j,j} = V^H W, W -= V H      block classical Gram-Schmidt, one pass
//                                                  as pgmres_iter.m:106-107 (or two)
//                 W = V_{j+1} H_{j+1,j}            shifted CholeskyQR3
//   least squares min || E_1 S_0 - Hbar Y ||  by Householder QR of the block
//   Hessenberg matrix, updated one block column per step on the host;
//   X += V Y; the true residual B - A X (FP64) starts the next cycle.
// The tall-skinny products (V^H W with K = n ~ 10^6 rows, and V C with a
// small C) are this file's kernels: a library GEMM tiles the small P x s
// output and leaves most of the chip idle on K; here k_vhw splits the rows
// over workgroups and runs its 16 x 16 output tiles on the matrix cores
// (operands staged in LDS), and k_vc streams V once per 16 output columns
// with the coefficients as uniform operands. The multi-RHS product is this
// library's FFT path, and may be replayed from HIP graphs captured per block
// step.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

namespace marie {

using cd = std::complex<double>;
constexpr int BT = 256;
constexpr int BG = 1024;

namespace {

template <typename TO, typename TI>
__global__ void k_convert(c2<TO>* __restrict__ o, const c2<TI>* __restrict__ in, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const c2<TI> v = in[e];
        o[e] = c2<TO>{(TO)v.x, (TO)v.y};
    }
}

// r = b - t
__global__ void k_sub(c2<double>* __restrict__ r, const c2<double>* __restrict__ b, const c2<double>* __restrict__ t,
                      int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        r[e] = c2<double>{b[e].x - t[e].x, b[e].y - t[e].y};
}

// x += d
__global__ void k_add(c2<double>* __restrict__ x, const c2<double>* __restrict__ d, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = c2<double>{x[e].x + d[e].x, x[e].y + d[e].y};
}

// part[col * G + blk] = block sum of |v_col|^2 (grid: G x s)
__global__ __launch_bounds__(BT) void k_colnorm2(const c2<double>* __restrict__ v, int64_t n,
                                                 double* __restrict__ part, int G) {
    __shared__ double sh[BT / 64];
    const c2<double>* col = v + (int64_t)blockIdx.y * n;
    double s = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x; e < n; e += (int64_t)G * BT)
        s += col[e].x * col[e].x + col[e].y * col[e].y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BT / 64; ++w) t += sh[w];
        part[(int64_t)blockIdx.y * G + blockIdx.x] = t;
    }
}

// out[col] = sum_b part[col * G + b], fixed order (one block per column)
__global__ __launch_bounds__(BT) void k_colsum(const double* __restrict__ part, int G, double* __restrict__ out) {
    __shared__ double sh[BT / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < G; b += BT) s += part[(int64_t)blockIdx.x * G + b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BT / 64; ++w) t += sh[w];
        out[blockIdx.x] = t;
    }
}

int grid_of(int64_t n) {
    int64_t g = (n + BT * 4 - 1) / (BT * 4);
    return (int)(g < 1 ? 1 : (g > BG ? BG : g));
}

// ---- tall-skinny products ----------------------------------------------------
// C (P x Q, partial per row block) = V^H W over rows, on the matrix cores
// (v_mfma_f64_16x16x4f64). V: P columns of length n (stride ldv), W: Q
// columns (stride ldw). A workgroup (4 waves) takes a tile of TP = 16 NSUB
// columns of V and the 16 columns q0.. of W (blockIdx.z), rows in chunks of
// KR taken rb, rb + G, ... (blockIdx.y). Each chunk of V and W is staged in
// LDS through loads of KR consecutive rows per column; each wave then runs
// the 4-row k-steps of its 16 x 16 output tile: with A = V^T (16 columns x 4
// rows) and B = W (4 rows x 16 columns) split into real and imaginary parts,
// conj(V)^T W = (Ar Br + Ai Bi) + j (Ar Bi - Ai Br): four real products on
// independent accumulators. When the tile has fewer columns than waves
// (WPT waves per 16-column subtile), the waves split the k-steps and add
// their sums in LDS. Partials go to part[(q P + p) G + rb].
//
// Operand layout of the instruction (16 x 4 A, 4 x 16 B, 16 x 16 D): lane
// i + 16 k holds A[i][k] and B[k][i]; D[r][c] is element r / 4 of lane
// c + 16 (r % 4).
constexpr int VQ = 16, KR = 32;
typedef double dx4 __attribute__((ext_vector_type(4)));

template <int WPT>
__global__ __launch_bounds__(BT) void k_vhw(const c2<double>* __restrict__ V, int64_t ldv, int P,
                                            const c2<double>* __restrict__ W, int64_t ldw, int Q, int64_t n,
                                            c2<double>* __restrict__ part, int G, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int NSUB = 4 / WPT, TP = 16 * NSUB;
    __shared__ c2<double> Vs[KR][TP + 1];
    __shared__ c2<double> Ws[KR][VQ + 1];
    __shared__ dx4 red[(WPT > 1) ? 4 * 64 * 2 : 1];
    const int p0 = blockIdx.x * TP;
    const int rb = blockIdx.y;
    const int q0 = blockIdx.z * VQ;
    const int pv = P - p0 < TP ? P - p0 : TP;
    const int qv = Q - q0 < VQ ? Q - q0 : VQ;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int psub = wave / WPT, kpart = wave % WPT;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t nchunk = (n + KR - 1) / KR;
    constexpr int LV = KR * TP / BT;                  // V elements staged per thread and chunk
    constexpr int LW = KR * VQ / BT;                  // W elements staged per thread and chunk
    c2<double> vreg[LV], wreg[LW];
    auto fetch = [&](int64_t ch) {
        const int64_t r0 = ch * KR;
#pragma unroll
        for (int k = 0; k < LV; ++k) {
            const int i = threadIdx.x + BT * k;
            const int r = i % KR, pc = i / KR;
            const int64_t e = r0 + r;
            vreg[k] = (pc < pv && e < n) ? V[(int64_t)(p0 + pc) * ldv + e] : c2<double>{0.0, 0.0};
        }
#pragma unroll
        for (int k = 0; k < LW; ++k) {
            const int i = threadIdx.x + BT * k;
            const int r = i % KR, qc = i / KR;
            const int64_t e = r0 + r;
            wreg[k] = (qc < qv && e < n) ? W[(int64_t)(q0 + qc) * ldw + e] : c2<double>{0.0, 0.0};
        }
    };
    dx4 crr = {0, 0, 0, 0}, cii = {0, 0, 0, 0}, cri = {0, 0, 0, 0}, cir = {0, 0, 0, 0};
    if (rb < nchunk) fetch(rb);
    for (int64_t ch = rb; ch < nchunk; ch += G) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < LV; ++k) {
            const int i = threadIdx.x + BT * k;
            Vs[i % KR][i / KR] = vreg[k];
        }
#pragma unroll
        for (int k = 0; k < LW; ++k) {
            const int i = threadIdx.x + BT * k;
            Ws[i % KR][i / KR] = wreg[k];
        }
        __syncthreads();
        if (ch + G < nchunk) fetch(ch + G);
#pragma unroll
        for (int ks = kpart; ks < KR / 4; ks += WPT) {
            const c2<double> a = Vs[4 * ks + lk][16 * psub + li];
            const c2<double> b = Ws[4 * ks + lk][li];
            crr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, crr, 0, 0, 0);
            cii = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, cii, 0, 0, 0);
            cri = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, cri, 0, 0, 0);
            cir = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, cir, 0, 0, 0);
        }
    }
    dx4 cr = crr + cii, ci = cri - cir;
    if constexpr (WPT > 1) {
        // the k-step partial sums of the waves of one subtile, added in a fixed order
        __syncthreads();
        red[(wave * 64 + lane) * 2] = cr;
        red[(wave * 64 + lane) * 2 + 1] = ci;
        __syncthreads();
        if (kpart != 0) return;
        for (int w2 = 1; w2 < WPT; ++w2) {
            cr += red[((wave + w2) * 64 + lane) * 2];
            ci += red[((wave + w2) * 64 + lane) * 2 + 1];
        }
    }
    // lane (c + 16 (r % 4)), element e: output row r = 4 e + lk (a column p of V), column li (a column q of W)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int pr = 16 * psub + 4 * e + lk;
        if (pr < pv && li < qv)
            part[((int64_t)(q0 + li) * P + p0 + pr) * G + rb] = c2<double>{cr[e], ci[e]};
    }
}

// C[q * ldc + p] = sum_rb part[(q P + p) G + rb]: one wave per output, lanes
// over rb in a fixed order, then a fixed shuffle tree
__global__ __launch_bounds__(BT) void k_vhw_reduce(const c2<double>* __restrict__ part, int P, int Q, int G,
                                                   c2<double>* __restrict__ C, int64_t ldc,
                                                   const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int o = blockIdx.x * (BT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= P * Q) return;
    const c2<double>* src = part + (int64_t)o * G;
    double x = 0.0, y = 0.0;
    for (int rb = lane; rb < G; rb += 64) {
        x += src[rb].x;
        y += src[rb].y;
    }
    for (int s = 32; s > 0; s >>= 1) {
        x += __shfl_down(x, s, 64);
        y += __shfl_down(y, s, 64);
    }
    if (lane == 0) {
        const int p = o % P, q = o / P;
        C[(int64_t)q * ldc + p] = c2<double>{x, y};
    }
}

// out(:, q) = beta * out(:, q) + alpha * sum_p V(:, p) C(p, q): thread per row,
// CQ output columns per workgroup. Every element of V meets CQ coefficients
// C(p, q0..q0+CQ-1), the same for all lanes: they are read as uniform
// (scalar) loads and enter the FMAs as SGPR operands, with no LDS in the loop.
constexpr int CQ = 16;
__global__ __launch_bounds__(BT) void k_vc(c2<double>* __restrict__ out, int64_t ldo, const c2<double>* __restrict__ V,
                                           int64_t ldv, int P, const c2<double>* __restrict__ C, int64_t ldc, int Q,
                                           int64_t n, double alpha, double beta, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int q0 = blockIdx.y * CQ;
    const int nq = Q - q0 < CQ ? Q - q0 : CQ;
    const int64_t e = blockIdx.x * (int64_t)BT + threadIdx.x;
    const bool live = e < n;
    const int64_t el = live ? e : 0;
    double ax[CQ] = {}, ay[CQ] = {};
    const c2<double>* cb = C + (int64_t)q0 * ldc;
    for (int p = 0; p < P; ++p) {
        const c2<double> v = V[(int64_t)p * ldv + el];
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
            // columns past nq repeat column nq-1; their sums are not stored
            const int qc = q < nq ? q : nq - 1;
            const c2<double> cc = cb[(int64_t)qc * ldc + p];
            ax[q] = fma(v.x, cc.x, fma(-v.y, cc.y, ax[q]));
            ay[q] = fma(v.x, cc.y, fma(v.y, cc.x, ay[q]));
        }
    }
    if (!live) return;
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
        if (q < nq) {
            c2<double>* o = out + (int64_t)(q0 + q) * ldo + e;
            double ox = alpha * ax[q], oy = alpha * ay[q];
            if (beta != 0.0) {
                ox += beta * o->x;
                oy += beta * o->y;
            }
            *o = c2<double>{ox, oy};
        }
    }
}

struct BlockWs {
    c2<double>* V = nullptr;     // n x (m+1)s
    c2<double>* W = nullptr;     // n x s
    c2<double>* T = nullptr;     // n x s (CholeskyQR staging, products)
    c2<double>* R = nullptr;     // n x s (true residual)
    c2<double>* D = nullptr;     // n x s (cycle correction)
    c2<float>* Xs = nullptr;     // n x s complex64 staging (mode 1)
    c2<float>* Ys = nullptr;
    c2<double>* Hd = nullptr;    // (m+1)s x s block column / Gram / small factors
    c2<double>* Sd = nullptr;    // (m+1)s x s small matrix upload
    double* part = nullptr;      // s x BG
    double* nrm = nullptr;       // s
    c2<double>* vpart = nullptr; // k_vhw partials (VPART complex)
    std::vector<hipGraphExec_t> graphs;
    std::vector<hipGraph_t> gsrc;
    void free_all() {
        void* p[] = {V, W, T, R, D, Xs, Ys, Hd, Sd, part, nrm, vpart};
        for (void* q : p)
            if (q) (void)hipFree(q);
        for (auto g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto g : gsrc)
            if (g) (void)hipGraphDestroy(g);
    }
};

struct BCtx {
    Op* h;
    hipStream_t s;
    int64_t n;
    int ns;                      // block size s
    int G;
    BlockWs ws;
};

// row blocks of k_vhw: about 1024 workgroups in all, 64..512 per output tile;
// the partials then take at most VPART complex values
constexpr size_t VPART = (size_t)1024 * 256 * VQ;
int vhw_blocks(int tiles) {
    int g = 1024 / tiles;
    return g < 64 ? 64 : (g > 512 ? 512 : g);
}

// C (P x Q, ld ldc) = V^H W over n rows (nothing when *skip is set on the device)
int vhw(BCtx& c, const c2<double>* V, int P, const c2<double>* W, int Q, c2<double>* C, int64_t ldc,
        const int* skip = nullptr) {
    // waves per 16-column subtile: a narrow V (the CholeskyQR Gram, early
    // block steps) splits the k-steps over the four waves instead
    const int wpt = P <= 16 ? 4 : (P <= 32 ? 2 : 1);
    const int tp = 16 * (4 / wpt);
    const int ptiles = (P + tp - 1) / tp, qtiles = (Q + VQ - 1) / VQ;
    const int G = vhw_blocks(ptiles * qtiles);
    if ((size_t)G * P * Q > VPART) {
        set_error("block too large for the multi-dot partials");
        return ERR_SIZE;
    }
    const dim3 grid(ptiles, G, qtiles);
    if (wpt == 4)
        hipLaunchKernelGGL(k_vhw<4>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G, skip);
    else if (wpt == 2)
        hipLaunchKernelGGL(k_vhw<2>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G, skip);
    else
        hipLaunchKernelGGL(k_vhw<1>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G, skip);
    hipLaunchKernelGGL(k_vhw_reduce, dim3((P * Q + BT / 64 - 1) / (BT / 64)), dim3(BT), 0, c.s, c.ws.vpart, P, Q, G,
                       C, ldc, skip);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// out (n x Q) = beta * out + alpha * V (n x P) C (P x Q, ld ldc)
int vc(BCtx& c, c2<double>* out, const c2<double>* V, int P, const c2<double>* C, int64_t ldc, int Q, double alpha,
       double beta, const int* skip = nullptr) {
    const dim3 g((unsigned)((c.n + BT - 1) / BT), (unsigned)((Q + CQ - 1) / CQ));
    hipLaunchKernelGGL(k_vc, g, dim3(BT), 0, c.s, out, c.n, V, c.n, P, C, ldc, Q, c.n, alpha, beta, skip);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// per-column 2-norms of an n x s block (host result)
int colnorms(BCtx& c, const c2<double>* v, std::vector<double>& out) {
    hipLaunchKernelGGL(k_colnorm2, dim3(c.G, c.ns), dim3(BT), 0, c.s, v, c.n, c.ws.part, c.G);
    hipLaunchKernelGGL(k_colsum, dim3(c.ns), dim3(BT), 0, c.s, c.ws.part, c.G, c.ws.nrm);
    MARIE_HIP(hipGetLastError());
    out.assign(c.ns, 0.0);
    MARIE_HIP(hipMemcpyAsync(out.data(), c.ws.nrm, c.ns * sizeof(double), hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipStreamSynchronize(c.s));
    for (double& x : out) x = std::sqrt(x);
    return OK;
}

// Hermitian positive definite s x s (column-major) -> upper R with A = R^H R;
// false when A is not numerically positive definite
bool cholesky_upper(const std::vector<cd>& A, int s, std::vector<cd>& R, double min_pivot = 1e-28) {
    R.assign((size_t)s * s, cd(0, 0));
    double amax = 0.0;
    for (int i = 0; i < s; ++i) amax = std::max(amax, A[(size_t)i * s + i].real());
    for (int j = 0; j < s; ++j) {
        cd d = A[(size_t)j * s + j];
        for (int k = 0; k < j; ++k) d -= std::conj(R[(size_t)j * s + k]) * R[(size_t)j * s + k];
        if (!(d.real() > min_pivot * amax) || !std::isfinite(d.real())) return false;
        const double rjj = std::sqrt(d.real());
        R[(size_t)j * s + j] = rjj;
        for (int i = j + 1; i < s; ++i) {
            cd v = A[(size_t)i * s + j];   // A(j, i)
            for (int k = 0; k < j; ++k) v -= std::conj(R[(size_t)j * s + k]) * R[(size_t)i * s + k];
            R[(size_t)i * s + j] = v / rjj;   // R(j, i)
        }
    }
    return true;
}

// inverse of an upper triangular s x s matrix (column-major)
std::vector<cd> upper_inverse(const std::vector<cd>& R, int s) {
    std::vector<cd> X((size_t)s * s, cd(0, 0));
    for (int j = 0; j < s; ++j) {
        X[(size_t)j * s + j] = 1.0 / R[(size_t)j * s + j];
        for (int i = j - 1; i >= 0; --i) {
            cd v = 0;
            for (int k = i + 1; k <= j; ++k) v += R[(size_t)k * s + i] * X[(size_t)j * s + k];
            X[(size_t)j * s + i] = -v / R[(size_t)i * s + i];
        }
    }
    return X;
}

// CholeskyQR: W = out Rf with out orthonormal (out != W). Each pass
// equilibrates the Gram matrix to unit diagonal first, so columns of very
// different norms (a nearly converged right-hand side next to others) lose
// nothing. When the first Gram matrix factors with every pivot above 1e-10
// (condition number below 1e5) two plain passes suffice (CholeskyQR2);
// otherwise the first pass factors the Gram matrix plus a small shift
// (11 (n s + s(s+1)) u per unit diagonal), which always succeeds and leaves a
// basis of condition number O(u^-1/2), and two plain passes follow (shifted
// CholeskyQR3). Returns ERR_STATE (no error text) for a zero or non-finite
// column.
int cholqr(BCtx& c, const c2<double>* W, c2<double>* out, std::vector<cd>& Rf) {
    const int s = c.ns;
    std::vector<cd> g((size_t)s * s);
    std::vector<std::vector<cd>> Rp(3);
    const double shift = 11.0 * ((double)c.n * s + (double)s * (s + 1)) * 1.1102230246251565e-16 * s;
    const c2<double>* src = W;
    c2<double>* dsts[3] = {out, c.ws.T, out};
    int npass = 3;
    for (int pass = 0; pass < npass; ++pass) {
        int st = vhw(c, src, s, src, s, c.ws.Hd, s);
        if (st) return st;
        MARIE_HIP(hipMemcpyAsync(g.data(), c.ws.Hd, (size_t)s * s * sizeof(c2<double>), hipMemcpyDeviceToHost, c.s));
        MARIE_HIP(hipStreamSynchronize(c.s));
        std::vector<double> d(s);
        for (int i = 0; i < s; ++i) {
            d[i] = std::sqrt(g[(size_t)i * s + i].real());
            if (!(d[i] > 0.0) || !std::isfinite(d[i])) return ERR_STATE;
        }
        for (int jj = 0; jj < s; ++jj)
            for (int i = 0; i < s; ++i) g[(size_t)jj * s + i] /= d[i] * d[jj];
        std::vector<cd>& R = Rp[pass];
        if (pass == 0 && cholesky_upper(g, s, R, 1e-10)) {
            npass = 2;                  // well conditioned: CholeskyQR2
            dsts[0] = c.ws.T;
            dsts[1] = out;
        } else {
            if (pass == 0)
                for (int i = 0; i < s; ++i) g[(size_t)i * s + i] += shift;
            if (!cholesky_upper(g, s, R)) return ERR_STATE;
        }
        for (int jj = 0; jj < s; ++jj)
            for (int i = 0; i <= jj; ++i) R[(size_t)jj * s + i] *= d[jj];
        const std::vector<cd> Ri = upper_inverse(R, s);
        MARIE_HIP(hipMemcpyAsync(c.ws.Sd, Ri.data(), (size_t)s * s * sizeof(c2<double>), hipMemcpyHostToDevice,
                                 c.s));
        st = vc(c, dsts[pass], src, s, c.ws.Sd, s, s, 1.0, 0.0);
        if (st) return st;
        MARIE_HIP(hipStreamSynchronize(c.s));   // Ri (host memory) goes out of scope
        src = dsts[pass];
    }
    // Rf = R3 R2 R1
    auto mul_upper = [s](const std::vector<cd>& A, const std::vector<cd>& B) {
        std::vector<cd> C((size_t)s * s, cd(0, 0));
        for (int j = 0; j < s; ++j)
            for (int i = 0; i <= j; ++i) {
                cd v = 0;
                for (int k = i; k <= j; ++k) v += A[(size_t)k * s + i] * B[(size_t)j * s + k];
                C[(size_t)j * s + i] = v;
            }
        return C;
    };
    Rf = npass == 2 ? mul_upper(Rp[1], Rp[0]) : mul_upper(Rp[2], mul_upper(Rp[1], Rp[0]));
    return OK;
}

// W = A V_j through the FFT path, in complex128 or (mode 1) complex64
int block_matvec(BCtx& c, const c2<double>* Vj, c2<double>* W, bool f32, int j, bool use_graph) {
    const int64_t ns = c.n * c.ns;
    auto body = [&]() -> int {
        if (!f32) return apply(c.h, OP_AN, 0, PREC_F64, Vj, W, c.ns, c.s);
        const int g = grid_of(ns);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(BT), 0, c.s, c.ws.Xs, Vj, ns);
        int st = apply(c.h, OP_AN, 0, PREC_F32, c.ws.Xs, c.ws.Ys, c.ns, c.s);
        if (st) return st;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(BT), 0, c.s, W, c.ws.Ys, ns);
        MARIE_HIP(hipGetLastError());
        return OK;
    };
    if (!use_graph || j < 0) return body();
    if ((int)c.ws.graphs.size() <= j) {
        c.ws.graphs.resize(j + 1, nullptr);
        c.ws.gsrc.resize(j + 1, nullptr);
    }
    if (!c.ws.graphs[j]) {
        MARIE_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
        const int st = body();
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c.s, &gr);
        if (st) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        if (e != hipSuccess) {
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            return ERR_HIP;
        }
        c.ws.gsrc[j] = gr;
        MARIE_HIP(hipGraphInstantiate(&c.ws.graphs[j], gr, nullptr, nullptr, 0));
    }
    MARIE_HIP(hipGraphLaunch(c.ws.graphs[j], c.s));
    return OK;
}

int alloc_block(BCtx& c, int m, int mode) {
    const int64_t n = c.n, s = c.ns;
    auto A = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the block Krylov workspace");
            return ERR_ALLOC;
        }
        return OK;
    };
    int st = A((void**)&c.ws.V, (size_t)(m + 1) * s * n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.W, (size_t)s * n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.T, (size_t)s * n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.R, (size_t)s * n * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.D, (size_t)s * n * sizeof(c2<double>));
    if (!st && mode == 1) st = A((void**)&c.ws.Xs, (size_t)s * n * sizeof(c2<float>));
    if (!st && mode == 1) st = A((void**)&c.ws.Ys, (size_t)s * n * sizeof(c2<float>));
    if (!st) st = A((void**)&c.ws.Hd, (size_t)(m + 1) * s * s * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.Sd, (size_t)(m + 1) * s * s * sizeof(c2<double>));
    if (!st) st = A((void**)&c.ws.part, (size_t)s * BG * sizeof(double));
    if (!st) st = A((void**)&c.ws.nrm, (size_t)s * sizeof(double));
    if (!st) st = A((void**)&c.ws.vpart, VPART * sizeof(c2<double>));
    return st;
}

// Householder QR of the block Hessenberg matrix, one block column at a time.
struct BlockLS {
    int s, m, ld;                      // ld = (m+1)s rows
    std::vector<cd> H;                 // ld x ms, column-major (original)
    std::vector<cd> Rq;                // transformed copy
    std::vector<cd> G;                 // ld x s right-hand side (transformed)
    std::vector<std::vector<cd>> hv;   // reflector of column c (length s+1), rows c..c+s
    std::vector<double> hb;            // 2 / (v^H v), 0 for an identity reflector
    void init(int s_, int m_, const std::vector<cd>& S0) {
        s = s_; m = m_; ld = (m + 1) * s;
        H.assign((size_t)ld * m * s, cd(0, 0));
        Rq = H;
        G.assign((size_t)ld * s, cd(0, 0));
        for (int q = 0; q < s; ++q)
            for (int i = 0; i <= q; ++i) G[(size_t)q * ld + i] = S0[(size_t)q * s + i];
        hv.assign((size_t)m * s, std::vector<cd>());
        hb.assign((size_t)m * s, 0.0);
    }
    void reflect(int c, cd* col) const {        // apply reflector c to a column of length ld
        if (hb[c] == 0.0) return;
        const std::vector<cd>& v = hv[c];
        cd d = 0;
        for (int i = 0; i <= s; ++i) d += std::conj(v[i]) * col[c + i];
        d *= hb[c];
        for (int i = 0; i <= s; ++i) col[c + i] -= v[i] * d;
    }
    // block column j: Hcol is ld x s (rows beyond (j+2)s zero); returns per-column LS residuals
    void add_block(int j, const std::vector<cd>& Hcol, std::vector<double>& res) {
        for (int q = 0; q < s; ++q) {
            const int c = j * s + q;
            cd* hc = &H[(size_t)c * ld];
            cd* rc = &Rq[(size_t)c * ld];
            for (int i = 0; i < ld; ++i) hc[i] = rc[i] = Hcol[(size_t)q * ld + i];
            for (int p = 0; p < c; ++p) reflect(p, rc);
            // new reflector zeroing rows c+1..c+s of this column
            std::vector<cd> v(s + 1);
            double nx = 0.0;
            for (int i = 0; i <= s; ++i) {
                v[i] = rc[c + i];
                nx += std::norm(v[i]);
            }
            nx = std::sqrt(nx);
            double tail = nx * nx - std::norm(v[0]);
            if (nx == 0.0 || tail <= 0.0) {
                hb[c] = 0.0;
            } else {
                const double a0 = std::abs(v[0]);
                const cd ph = a0 > 0 ? v[0] / a0 : cd(1, 0);
                v[0] += ph * nx;
                double vv = 0.0;
                for (int i = 0; i <= s; ++i) vv += std::norm(v[i]);
                hv[c] = v;
                hb[c] = 2.0 / vv;
                reflect(c, rc);
                for (int i = 1; i <= s; ++i) rc[c + i] = 0;
                for (int q2 = 0; q2 < s; ++q2) reflect(c, &G[(size_t)q2 * ld]);
            }
        }
        res.assign(s, 0.0);
        for (int q = 0; q < s; ++q) {
            double r2 = 0.0;
            for (int i = (j + 1) * s; i < (j + 2) * s; ++i) r2 += std::norm(G[(size_t)q * ld + i]);
            res[q] = std::sqrt(r2);
        }
    }
    // Y (Js x s, column-major) = R^-1 G(0:Js, :)
    std::vector<cd> solve(int J) const {
        const int K = J * s;
        std::vector<cd> Y((size_t)K * s, cd(0, 0));
        for (int q = 0; q < s; ++q)
            for (int i = K - 1; i >= 0; --i) {
                cd v = G[(size_t)q * ld + i];
                for (int k = i + 1; k < K; ++k) v -= Rq[(size_t)k * ld + i] * Y[(size_t)q * K + k];
                Y[(size_t)q * K + i] = v / Rq[(size_t)i * ld + i];
            }
        return Y;
    }
};

// One cycle on the current residual R: D = V Y with the block Krylov space of R.
// Returns the number of block steps, or -1 with *st set on error.
int block_cycle(BCtx& c, int m, const std::vector<double>& target, bool f32, bool use_graph, int ncgs, int* steps,
                std::vector<double>& hist, const std::vector<double>& bn) {
    const int s = c.ns;
    const int64_t n = c.n;
    std::vector<cd> S0;
    int st = cholqr(c, c.ws.R, c.ws.V, S0);
    if (st == ERR_STATE) {
        set_error("the block of residuals is rank deficient (dependent right-hand sides)");
        return ERR_ARG;
    }
    if (st) return st;
    BlockLS ls;
    ls.init(s, m, S0);
    const int ld = (m + 1) * s;
    std::vector<cd> Hcol((size_t)ld * s), Ht((size_t)ld * s);
    std::vector<double> res;
    int J = 0;
    for (int j = 0; j < m; ++j) {
        c2<double>* Vj = c.ws.V + (int64_t)j * s * n;
        st = block_matvec(c, Vj, c.ws.W, f32, j, use_graph);
        if (st) return st;
        const int P = (j + 1) * s;
        std::fill(Hcol.begin(), Hcol.end(), cd(0, 0));
        for (int pass = 0; pass < ncgs; ++pass) {   // block classical Gram-Schmidt passes
            st = vhw(c, c.ws.V, P, c.ws.W, s, c.ws.Hd, P);
            if (st) return st;
            st = vc(c, c.ws.W, c.ws.V, P, c.ws.Hd, P, s, -1.0, 1.0);
            if (st) return st;
            MARIE_HIP(hipMemcpyAsync(Ht.data(), c.ws.Hd, (size_t)P * s * sizeof(c2<double>), hipMemcpyDeviceToHost,
                                     c.s));
            MARIE_HIP(hipStreamSynchronize(c.s));
            for (int q = 0; q < s; ++q)
                for (int i = 0; i < P; ++i) Hcol[(size_t)q * ld + i] += Ht[(size_t)q * P + i];
        }
        std::vector<cd> Rs;
        st = cholqr(c, c.ws.W, c.ws.V + (int64_t)(j + 1) * s * n, Rs);
        const bool breakdown = (st == ERR_STATE);
        if (st && !breakdown) return st;
        if (!breakdown)
            for (int q = 0; q < s; ++q)
                for (int i = 0; i <= q; ++i) Hcol[(size_t)q * ld + P + i] = Rs[(size_t)q * s + i];
        ls.add_block(j, Hcol, res);
        J = j + 1;
        double worst = 0.0;
        bool done = true;
        for (int q = 0; q < s; ++q) {
            worst = std::max(worst, res[q] / bn[q]);
            if (res[q] > target[q]) done = false;
        }
        hist.push_back(worst);
        if (done || breakdown) break;
    }
    // D = V(:, 0:Js) Y
    const std::vector<cd> Y = ls.solve(J);
    MARIE_HIP(hipMemcpyAsync(c.ws.Sd, Y.data(), Y.size() * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    st = vc(c, c.ws.D, c.ws.V, J * s, c.ws.Sd, (int64_t)J * s, s, 1.0, 0.0);
    if (st) return st;
    MARIE_HIP(hipStreamSynchronize(c.s));
    *steps = J;
    return OK;
}

// ---- device-driven block steps -----------------------------------------------
// Every part of a block step runs on the device: the product, the block
// Gram-Schmidt pass, CholeskyQR (k_bchol: the s x s factor on the device, the
// third pass skipped on the device when the first factored unshifted) and the
// Householder update of the least-squares problem (k_bls). A step is then one
// stream-ordered sequence with no host round trip, captured once per step
// index into a HIP graph and replayed in every cycle; the host reads one
// status record per step (the columns' least-squares residuals) and decides
// whether to take another step or close the cycle (k_bsolve, V Y, the true
// residual). One host thread drives any number of such solves by their
// events (marie_vie_block_gmres_batch).

constexpr int DS = 32;           // most right-hand sides of a device-driven block
constexpr int DLS = 1024;        // threads of the least-squares update kernel

struct DevStatus {
    double res[DS];              // |LS residual| per column after the last block step
    double rn2[DS];              // squared true residual norms (cycle end)
    int skip3;                   // CholeskyQR: the first pass factored unshifted, the third is skipped
    int brk;                     // CholeskyQR breakdown in the last block step
    int brk0;                    // breakdown of the cycle's first CholeskyQR (the residual block)
    int pad;
};

__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 dcmul(double2 a, double2 b) {      // conj(a) b
    return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double2 ld2d(const c2<double>* p) {
    const c2<double> v = *p;
    return double2{v.x, v.y};
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One CholeskyQR factor on the device (s <= DS): the Gram Gm (s x s,
// column-major) equilibrated to unit diagonal and factored R^H R (R upper,
// real positive diagonal; right-looking, one thread per trailing entry).
// Pass 0 first tries the plain factor with pivots above 1e-10 of the largest
// diagonal entry (then *skip3 = 1: two passes suffice) and otherwise factors
// Gram + shift I (*skip3 = 0); later passes factor plainly (pivots above
// 1e-28). R is scaled back (R D), Rinv = R^-1 (column-major, the k_vc
// coefficient layout), Rf = R (pass 0) or R Rf. A zero, non-finite or
// indefinite Gram sets *brk and zeroes Rinv and Rf; later passes of the same
// factorisation do nothing once *brk is set (pass 0 clears it).
__global__ __launch_bounds__(BT) void k_bchol(const c2<double>* __restrict__ Gm, int s, int pass, double shift,
                                              c2<double>* __restrict__ Rinv, c2<double>* __restrict__ Rf,
                                              int* __restrict__ skip3, int* __restrict__ brk,
                                              const int* __restrict__ skip) {
    if (skip && *skip) return;
    if (pass > 0 && *brk) return;
    __shared__ double2 A[DS][DS + 1];     // the equilibrated Gram, then R (upper)
    __shared__ double2 X[DS][DS + 1];     // R^-1
    __shared__ double2 F[DS][DS + 1];     // the previous Rf
    __shared__ double d[DS], rinv[DS];
    __shared__ double amax;
    __shared__ int ok, bad;
    const int t = threadIdx.x;
    if (t == 0) bad = 0;
    __syncthreads();
    if (t < s) {
        const double dd = sqrt(Gm[(int64_t)t * s + t].x);
        d[t] = dd;
        if (!(dd > 0.0) || !isfinite(dd)) bad = 1;
    }
    if (pass > 0)
        for (int e = t; e < DS * DS; e += BT) {
            const int r = e % DS, cc = e / DS;
            F[r][cc] = (r < s && cc < s) ? ld2d(Rf + (int64_t)cc * s + r) : double2{0.0, 0.0};
        }
    __syncthreads();
    bool factored = false;
    if (!bad) {
        for (int attempt = 0;; ++attempt) {
            const double minp = (pass == 0 && attempt == 0) ? 1e-10 : 1e-28;
            for (int e = t; e < DS * DS; e += BT) {
                const int r = e % DS, cc = e / DS;
                double2 v = {0.0, 0.0};
                if (r < s && cc < s) {
                    const c2<double> g = Gm[(int64_t)cc * s + r];
                    const double sc = 1.0 / (d[r] * d[cc]);
                    v = double2{g.x * sc, g.y * sc};
                    if (r == cc && attempt == 1) v.x += shift;
                }
                A[r][cc] = v;
                X[r][cc] = double2{0.0, 0.0};
            }
            __syncthreads();
            if (t == 0) {
                double am = 0.0;
                for (int i = 0; i < s; ++i) am = fmax(am, A[i][i].x);
                amax = am;
                ok = 0;
            }
            __syncthreads();
            for (int jj = 0; jj < s; ++jj) {
                if (t == 0) {
                    const double piv = A[jj][jj].x;
                    if (piv > minp * amax && isfinite(piv)) {
                        const double rjj = sqrt(piv);
                        A[jj][jj] = double2{rjj, 0.0};
                        rinv[jj] = 1.0 / rjj;
                        ok = jj + 1;
                    }
                }
                __syncthreads();
                if (ok != jj + 1) break;
                if (t > jj && t < s) A[jj][t] = double2{A[jj][t].x * rinv[jj], A[jj][t].y * rinv[jj]};
                __syncthreads();
                for (int e = t; e < DS * DS; e += BT) {
                    const int a = e / DS, b = e % DS;
                    if (a > jj && b >= a && b < s) {
                        const double2 p = dcmul(A[jj][a], A[jj][b]);
                        A[a][b].x -= p.x;
                        A[a][b].y -= p.y;
                    }
                }
                __syncthreads();
            }
            if (ok == s) {
                factored = true;
                if (pass == 0 && t == 0) *skip3 = attempt == 0 ? 1 : 0;
                break;
            }
            if (pass != 0 || attempt == 1) break;
            __syncthreads();
        }
    }
    if (!factored) {
        for (int e = t; e < s * s; e += BT) {
            Rinv[e] = c2<double>{0.0, 0.0};
            Rf[e] = c2<double>{0.0, 0.0};
        }
        if (t == 0) {
            *brk = 1;
            if (pass == 0) *skip3 = 1;
        }
        return;
    }
    if (pass == 0 && t == 0) *brk = 0;
    // R D, then R^-1 column by column
    for (int e = t; e < DS * DS; e += BT) {
        const int r = e % DS, cc = e / DS;
        if (r <= cc && cc < s) A[r][cc] = double2{A[r][cc].x * d[cc], A[r][cc].y * d[cc]};
    }
    __syncthreads();
    if (t < s) rinv[t] = 1.0 / A[t][t].x;
    __syncthreads();
    if (t < s) {
        const int l = t;
        X[l][l] = double2{rinv[l], 0.0};
        for (int i = l - 1; i >= 0; --i) {
            double2 v = {0.0, 0.0};
            for (int k = i + 1; k <= l; ++k) {
                const double2 p = dmul(A[i][k], X[k][l]);
                v.x += p.x;
                v.y += p.y;
            }
            X[i][l] = double2{-v.x * rinv[i], -v.y * rinv[i]};
        }
    }
    __syncthreads();
    for (int e = t; e < s * s; e += BT) {
        const int r = e % s, cc = e / s;
        const double2 xi = r <= cc ? X[r][cc] : double2{0.0, 0.0};
        Rinv[e] = c2<double>{xi.x, xi.y};
        double2 f = {0.0, 0.0};
        if (r <= cc) {
            if (pass == 0) {
                f = A[r][cc];
            } else {
                for (int k = r; k <= cc; ++k) {
                    const double2 p = dmul(A[r][k], F[k][cc]);
                    f.x += p.x;
                    f.y += p.y;
                }
            }
        }
        Rf[e] = c2<double>{f.x, f.y};
    }
}

// Least-squares update of block step j (one workgroup, s <= DS): the new block
// column (columns js .. js+s-1 of Rq, ld rows, nonzero in rows 0 .. (j+2)s-1)
// meets the reflectors of the earlier columns in order, then its own:
// reflector c zeroes rows c+1 .. c+s of column c (v = x + phase(x_0) ||x|| e_0,
// beta = 2 / v^H v, the identity when x is already zero below row c) and is
// applied to the block's later columns and to the right-hand side G (ld x s).
// res[q] = || G(rows (j+1)s .. (j+2)s-1, column q) ||.
// The s reflectors of block b act on rows bs .. bs+2s-1 only; their product
// H_0 .. H_{s-1} = I - V T V^H (V: 2s x s, column i holding reflector bs+i at
// rows i .. i+s; T upper triangular, T_ii = beta_i, T(0:i, i) = -beta_i
// T(0:i, 0:i) V(:, 0:i)^H v_i) is kept per block, so the earlier blocks reach
// the new columns as one 2s-row window update each, W -= V (T^H (V^H W)),
// three small products in LDS; rows bs .. bs+s-1 are then final and the
// window moves down by s rows. Block j's own reflectors work on the window
// rows js .. js+2s-1 one after the other, and block j's T is formed beside
// them.
__global__ __launch_bounds__(DLS) void k_bls(c2<double>* __restrict__ Rq, int64_t ld, c2<double>* __restrict__ G,
                                             c2<double>* __restrict__ hv, double* __restrict__ hb,
                                             c2<double>* __restrict__ Tb, int s, int j, double* __restrict__ res) {
    constexpr int NW = DLS / 64;
    __shared__ double2 Wn[2 * DS][DS + 1];   // the window of the new columns
    __shared__ double2 Gw[2 * DS][DS + 1];   // G's window (phase B); V^H W (phase A)
    __shared__ double2 Vs[2 * DS][DS + 1];   // a block's reflectors as V
    __shared__ double2 Ts[DS][DS + 1];       // a block's T
    __shared__ double2 C2[DS][DS + 1];       // T^H V^H W
    __shared__ double2 vcur[DS + 1];
    __shared__ double bcur;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const int js = j * s, L = s + 1, s2 = 2 * s;
    const double2 z = {0.0, 0.0};
    auto rq = [&](int row, int q) -> c2<double>* { return Rq + (int64_t)(js + q) * ld + row; };
    // the first window: rows 0 .. 2s-1 of the new columns
    for (int e = t; e < s2 * s; e += DLS) {
        const int r = e % s2, q = e / s2;
        Wn[r][q] = ld2d(rq(r, q));
    }
    // ---- earlier blocks b < j
    for (int b = 0; b < j; ++b) {
        const int bs = b * s;
        __syncthreads();
        for (int e = t; e < s2 * s; e += DLS) {
            const int r = e % s2, i = e / s2;
            const int k = r - i;
            Vs[r][i] = (k >= 0 && k <= s) ? ld2d(hv + (int64_t)(bs + i) * L + k) : z;
        }
        for (int e = t; e < s * s; e += DLS) {
            const int k = e % s, i = e / s;
            Ts[k][i] = ld2d(Tb + ((int64_t)b * s + i) * s + k);          // T(k, i), column-major per block
        }
        __syncthreads();
        // V^H W
        for (int e = t; e < s * s; e += DLS) {
            const int i = e % s, q = e / s;
            double2 a = z;
            for (int r = i; r <= i + s; ++r) {
                const double2 p = dcmul(Vs[r][i], Wn[r][q]);
                a.x += p.x;
                a.y += p.y;
            }
            Gw[i][q] = a;
        }
        __syncthreads();
        // T^H (V^H W)
        for (int e = t; e < s * s; e += DLS) {
            const int i = e % s, q = e / s;
            double2 a = z;
            for (int k = 0; k <= i; ++k) {
                const double2 p = dcmul(Ts[k][i], Gw[k][q]);
                a.x += p.x;
                a.y += p.y;
            }
            C2[i][q] = a;
        }
        __syncthreads();
        // W -= V C2
        for (int e = t; e < s2 * s; e += DLS) {
            const int r = e % s2, q = e / s2;
            double2 a = Wn[r][q];
            const int i0 = r - s > 0 ? r - s : 0, i1 = r < s - 1 ? r : s - 1;
            for (int i = i0; i <= i1; ++i) {
                const double2 p = dmul(Vs[r][i], C2[i][q]);
                a.x -= p.x;
                a.y -= p.y;
            }
            Wn[r][q] = a;
        }
        __syncthreads();
        // rows bs .. bs+s-1 are final; the window moves down by s rows
        for (int e = t; e < s * s; e += DLS) {
            const int r = e % s, q = e / s;
            const double2 v = Wn[r][q];
            *rq(bs + r, q) = c2<double>{v.x, v.y};
        }
        __syncthreads();
        for (int e = t; e < s * s; e += DLS) {
            const int r = e % s, q = e / s;
            Wn[r][q] = Wn[s + r][q];
            Wn[s + r][q] = ld2d(rq(bs + s2 + r, q));
        }
    }
    // ---- the block's own reflectors on rows js .. js+2s-1
    for (int e = t; e < s2 * s; e += DLS) {
        const int r = e % s2, q = e / s2;
        Gw[r][q] = ld2d(G + (int64_t)q * ld + js + r);
        Vs[r][q] = z;
    }
    for (int e = t; e < s * s; e += DLS) Ts[e % s][e / s] = z;
    __syncthreads();
    for (int q = 0; q < s; ++q) {
        const int c = js + q;
        if (wave == 0) {
            const double2 xv = lane < L ? Wn[q + lane][q] : z;
            const double nx2 = wave_sum(xv.x * xv.x + xv.y * xv.y);
            const double2 x0 = {__shfl(xv.x, 0, 64), __shfl(xv.y, 0, 64)};
            const double nx = sqrt(nx2);
            const double tail = nx2 - (x0.x * x0.x + x0.y * x0.y);
            double beta = 0.0;
            double2 v = z;
            if (!(nx == 0.0 || tail <= 0.0)) {
                const double a0 = sqrt(x0.x * x0.x + x0.y * x0.y);
                const double2 ph = a0 > 0.0 ? double2{x0.x / a0, x0.y / a0} : double2{1.0, 0.0};
                v = xv;
                if (lane == 0) {
                    v.x += ph.x * nx;
                    v.y += ph.y * nx;
                }
                beta = 2.0 / wave_sum(v.x * v.x + v.y * v.y);
                const double2 pr = dcmul(v, xv);
                const double dx = wave_sum(pr.x) * beta, dy = wave_sum(pr.y) * beta;
                const double2 dv = dmul(v, double2{dx, dy});
                if (lane < L) Wn[q + lane][q] = lane == 0 ? double2{xv.x - dv.x, xv.y - dv.y} : z;
            }
            if (lane < L) {
                vcur[lane] = v;
                Vs[q + lane][q] = v;
                hv[(int64_t)c * L + lane] = c2<double>{v.x, v.y};
            }
            if (lane == 0) {
                bcur = beta;
                hb[c] = beta;
                Ts[q][q] = double2{beta, 0.0};
            }
        }
        __syncthreads();
        const double beta = bcur;
        if (beta != 0.0) {
            const int nw = s - q - 1;               // the block's later columns, then G's s columns
            for (int k = wave; k < nw + s; k += NW) {
                const bool inW = k < nw;
                const int col = inW ? q + 1 + k : k - nw;
                const double2 xv = lane < L ? (inW ? Wn[q + lane][col] : Gw[q + lane][col]) : z;
                const double2 v = lane < L ? vcur[lane] : z;
                const double2 pr = dcmul(v, xv);
                const double dx = wave_sum(pr.x) * beta, dy = wave_sum(pr.y) * beta;
                const double2 dv = dmul(v, double2{dx, dy});
                if (lane < L) {
                    const double2 o = {xv.x - dv.x, xv.y - dv.y};
                    if (inW) Wn[q + lane][col] = o;
                    else Gw[q + lane][col] = o;
                }
            }
            // column q of T: C2[i][0] = v_i^H v_q (i < q), then T(0:q, q)
            if (t < q) {
                double2 a = z;
                for (int r = q; r <= t + s; ++r) {
                    const double2 p = dcmul(Vs[r][t], Vs[r][q]);
                    a.x += p.x;
                    a.y += p.y;
                }
                C2[t][0] = a;
            }
        }
        __syncthreads();
        if (beta != 0.0 && t < q) {
            double2 a = z;
            for (int k = t; k < q; ++k) {
                const double2 p = dmul(Ts[t][k], C2[k][0]);
                a.x += p.x;
                a.y += p.y;
            }
            Ts[t][q] = double2{-beta * a.x, -beta * a.y};
        }
        __syncthreads();
    }
    for (int e = t; e < s2 * s; e += DLS) {
        const int r = e % s2, q = e / s2;
        *rq(js + r, q) = c2<double>{Wn[r][q].x, Wn[r][q].y};
        G[(int64_t)q * ld + js + r] = c2<double>{Gw[r][q].x, Gw[r][q].y};
    }
    for (int e = t; e < s * s; e += DLS) {
        const int k = e % s, i = e / s;
        Tb[((int64_t)j * s + i) * s + k] = c2<double>{Ts[k][i].x, Ts[k][i].y};
    }
    if (t < s) {
        double r2 = 0.0;
        for (int r = s; r < s2; ++r) r2 += Gw[r][t].x * Gw[r][t].x + Gw[r][t].y * Gw[r][t].y;
        res[t] = sqrt(r2);
    }
}

// Y (K x s, column-major) = R^-1 G(0:K, :) with R the leading K x K upper
// triangle of Rq: one workgroup per column of G, column-oriented back
// substitution in LDS (y_k = g_k / R_kk, then g(0:k) -= R(0:k, k) y_k), the
// next column of R loaded while the current one is applied.
constexpr int BSR = 8;           // rows of R per thread: K <= BSR * BT
__global__ __launch_bounds__(BT) void k_bsolve(const c2<double>* __restrict__ Rq, int64_t ld,
                                               const c2<double>* __restrict__ G, int K, c2<double>* __restrict__ Y) {
    extern __shared__ double2 sm_[];
    double2* g = sm_;
    double2* dg = sm_ + K;
    const int q = blockIdx.x, t = threadIdx.x;
    for (int i = t; i < K; i += BT) {
        g[i] = ld2d(G + (int64_t)q * ld + i);
        dg[i] = ld2d(Rq + (int64_t)i * ld + i);
    }
    double2 cur[BSR], nxt[BSR];
#pragma unroll
    for (int i = 0; i < BSR; ++i) {
        const int r = t + BT * i;
        cur[i] = r < K - 1 ? ld2d(Rq + (int64_t)(K - 1) * ld + r) : double2{0.0, 0.0};
    }
    __syncthreads();
    for (int k = K - 1; k >= 0; --k) {
#pragma unroll
        for (int i = 0; i < BSR; ++i) {
            const int r = t + BT * i;
            nxt[i] = (k > 0 && r < k - 1) ? ld2d(Rq + (int64_t)(k - 1) * ld + r) : double2{0.0, 0.0};
        }
        const double2 gk = g[k], rk = dg[k];
        const double den = rk.x * rk.x + rk.y * rk.y;
        const double2 yk = {(gk.x * rk.x + gk.y * rk.y) / den, (gk.y * rk.x - gk.x * rk.y) / den};
#pragma unroll
        for (int i = 0; i < BSR; ++i) {
            const int r = t + BT * i;
            if (r < k) {
                const double2 p = dmul(cur[i], yk);
                g[r].x -= p.x;
                g[r].y -= p.y;
            }
        }
        if (t == 0) Y[(int64_t)q * K + k] = c2<double>{yk.x, yk.y};
        __syncthreads();
#pragma unroll
        for (int i = 0; i < BSR; ++i) cur[i] = nxt[i];
    }
}

// Rq block column col0 .. col0+s-1: rows 0 .. P-1 = Hd (P x s, column-major),
// or + Hd when acc; without acc the rows from P on are zeroed
__global__ void k_hset(const c2<double>* __restrict__ Hd, int P, int s, c2<double>* __restrict__ Rq, int64_t ld,
                       int col0, int acc) {
    const int64_t rows = acc ? P : ld;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < rows * s;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % rows;
        const int q = (int)(e / rows);
        c2<double>* o = Rq + (int64_t)(col0 + q) * ld + r;
        if (r < P) {
            const c2<double> h = Hd[(int64_t)q * P + r];
            *o = acc ? c2<double>{o->x + h.x, o->y + h.y} : h;
        } else {
            *o = c2<double>{0.0, 0.0};
        }
    }
}

// rows P .. P+s-1 of the block column = Rf (upper, column-major s x s)
__global__ void k_rset(const c2<double>* __restrict__ Rf, int s, c2<double>* __restrict__ Rq, int64_t ld, int col0,
                       int P) {
    const int e = threadIdx.x;
    if (e >= s * s) return;
    const int i = e % s, q = e / s;
    Rq[(int64_t)(col0 + q) * ld + P + i] = i <= q ? Rf[(int64_t)q * s + i] : c2<double>{0.0, 0.0};
}

// G (ld x s) = [S0; 0] with S0 upper (column-major s x s)
__global__ void k_ginit(const c2<double>* __restrict__ S0, int s, c2<double>* __restrict__ G, int64_t ld) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ld * s; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % ld;
        const int q = (int)(e / ld);
        G[e] = (i <= q) ? S0[(int64_t)q * s + i] : c2<double>{0.0, 0.0};
    }
}

// dst = src, nothing when *skip
__global__ void k_copy_skip(c2<double>* __restrict__ dst, const c2<double>* __restrict__ src, int64_t n,
                            const int* __restrict__ skip) {
    if (*skip) return;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

struct DevWs {
    c2<double>* Rq = nullptr;    // ld x m s transformed block Hessenberg
    c2<double>* Gd = nullptr;    // ld x s transformed right-hand side
    c2<double>* hv = nullptr;    // m s reflectors of length s + 1
    double* hb = nullptr;        // m s reflector factors
    c2<double>* Tb = nullptr;    // m blocks of s x s: each block's reflectors as I - V T V^H
    c2<double>* Gm = nullptr;    // s x s CholeskyQR Gram
    c2<double>* Ri = nullptr;    // 3 s x s inverse factors
    c2<double>* Rf = nullptr;    // s x s accumulated factor
    DevStatus* st = nullptr;     // device status
    DevStatus* hs = nullptr;     // pinned mirror
    hipEvent_t ev = nullptr;
    std::vector<hipGraphExec_t> graphs;
    std::vector<hipGraph_t> gsrc;
    void free_all() {
        void* p[] = {Rq, Gd, hv, hb, Tb, Gm, Ri, Rf, st};
        for (void* q : p)
            if (q) (void)hipFree(q);
        if (hs) (void)hipHostFree(hs);
        if (ev) (void)hipEventDestroy(ev);
        for (auto g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto g : gsrc)
            if (g) (void)hipGraphDestroy(g);
    }
};

// One restarted block GMRES solve driven by events: every call of advance()
// after its event completed reads the status and enqueues the next step, the
// cycle's end, the next cycle or nothing (done). The arithmetic is that of
// block_gmres/block_cycle (same CholeskyQR passes, Householder updates and
// stopping rule); only where it runs differs.
struct DevSolve {
    Op* h = nullptr;
    BCtx c;
    DevWs d;
    const c2<double>* B = nullptr;
    c2<double>* X = nullptr;
    int s = 0, m = 0, mode = 0, maxit = 0, ncgs = 1;
    bool use_graph = false, f32 = false;
    double tol = 0.0, inner_tol = 0.0;
    std::vector<double> bn, rn, target, hist;
    int cycles = 0, j = 0, steps = 0;
    enum { IDLE, STEP, CYCLE_END, DONE } phase = IDLE;
    int err = OK;
    std::string msg;
    bool started = false;        // init() ran: the stream and workspace are the handle's
    int64_t ld() const { return (int64_t)(m + 1) * s; }

    DevSolve() {
        c.h = nullptr;
        c.s = nullptr;
        c.n = 0;
        c.ns = 0;
        c.G = 0;
    }
    ~DevSolve() {
        if (started) {
            if (c.s) (void)hipStreamSynchronize(c.s);
            else (void)hipDeviceSynchronize();
        }
        d.free_all();
        c.ws.free_all();
    }
    int fail(int code) {
        err = code;
        msg = get_error();
        phase = DONE;
        return code;
    }
    double worst_rel() const {
        double w = 0.0;
        for (int q = 0; q < s; ++q) w = std::max(w, rn[q] / bn[q]);
        return w;
    }
    int alloc() {
        int st = alloc_block(c, m, mode);
        if (st) return st;
        auto A = [&](void** p, size_t bytes) -> int {
            if (hipMalloc(p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                *p = nullptr;
                set_error("out of device memory for the block Krylov workspace");
                return ERR_ALLOC;
            }
            return OK;
        };
        const size_t L = (size_t)ld();
        if (!st) st = A((void**)&d.Rq, L * m * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.Gd, L * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.hv, (size_t)m * s * (s + 1) * sizeof(c2<double>));
        if (!st) st = A((void**)&d.hb, (size_t)m * s * sizeof(double));
        if (!st) st = A((void**)&d.Tb, (size_t)m * s * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.Gm, (size_t)s * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.Ri, (size_t)3 * s * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.Rf, (size_t)s * s * sizeof(c2<double>));
        if (!st) st = A((void**)&d.st, sizeof(DevStatus));
        if (st) return st;
        MARIE_HIP(hipMemset(d.Ri, 0, (size_t)3 * s * s * sizeof(c2<double>)));
        MARIE_HIP(hipMemset(d.st, 0, sizeof(DevStatus)));
        MARIE_HIP(hipHostMalloc((void**)&d.hs, sizeof(DevStatus), hipHostMallocDefault));
        MARIE_HIP(hipEventCreateWithFlags(&d.ev, hipEventDisableTiming));
        return OK;
    }
    // CholeskyQR of the n x s block W into out (orthonormal), factor in d.Rf;
    // *brk (device) set on breakdown
    int cholqr_dev(const c2<double>* W, c2<double>* out, int* brk) {
        const double shift = 11.0 * ((double)c.n * s + (double)s * (s + 1)) * 1.1102230246251565e-16 * s;
        c2<double>* T = c.ws.T;
        c2<double>* Ri[3] = {d.Ri, d.Ri + s * s, d.Ri + 2 * s * s};
        int* sk = &d.st->skip3;
        int st = vhw(c, W, s, W, s, d.Gm, s);
        if (st) return st;
        hipLaunchKernelGGL(k_bchol, dim3(1), dim3(BT), 0, c.s, d.Gm, s, 0, shift, Ri[0], d.Rf, sk, brk,
                           (const int*)nullptr);
        st = vc(c, T, W, s, Ri[0], s, s, 1.0, 0.0);
        if (!st) st = vhw(c, T, s, T, s, d.Gm, s);
        if (st) return st;
        hipLaunchKernelGGL(k_bchol, dim3(1), dim3(BT), 0, c.s, d.Gm, s, 1, shift, Ri[1], d.Rf, sk, brk,
                           (const int*)nullptr);
        st = vc(c, out, T, s, Ri[1], s, s, 1.0, 0.0);
        // third pass only after a shifted first factor
        if (!st) st = vhw(c, out, s, out, s, d.Gm, s, sk);
        if (st) return st;
        hipLaunchKernelGGL(k_bchol, dim3(1), dim3(BT), 0, c.s, d.Gm, s, 2, shift, Ri[2], d.Rf, sk, brk,
                           (const int*)sk);
        st = vc(c, T, out, s, Ri[2], s, s, 1.0, 0.0, sk);
        if (st) return st;
        hipLaunchKernelGGL(k_copy_skip, dim3(grid_of(c.n * s)), dim3(BT), 0, c.s, out, T, c.n * s, (const int*)sk);
        MARIE_HIP(hipGetLastError());
        return OK;
    }
    // block step jj: W = A V_jj, Gram-Schmidt, CholeskyQR into V_jj+1, LS update
    int step_body(int jj) {
        const int64_t n = c.n;
        const int P = (jj + 1) * s;
        int st = block_matvec(c, c.ws.V + (int64_t)jj * s * n, c.ws.W, f32, -1, false);
        if (st) return st;
        for (int pass = 0; pass < ncgs; ++pass) {
            st = vhw(c, c.ws.V, P, c.ws.W, s, c.ws.Hd, P);
            if (!st) st = vc(c, c.ws.W, c.ws.V, P, c.ws.Hd, P, s, -1.0, 1.0);
            if (st) return st;
            const int64_t rows = pass > 0 ? P : ld();
            const int g = (int)std::min<int64_t>(BG, (rows * s + BT - 1) / BT);
            hipLaunchKernelGGL(k_hset, dim3(g), dim3(BT), 0, c.s, c.ws.Hd, P, s, d.Rq, ld(), jj * s, pass > 0 ? 1 : 0);
        }
        st = cholqr_dev(c.ws.W, c.ws.V + (int64_t)(jj + 1) * s * n, &d.st->brk);
        if (st) return st;
        hipLaunchKernelGGL(k_rset, dim3(1), dim3(DS * DS), 0, c.s, d.Rf, s, d.Rq, ld(), jj * s, P);
        hipLaunchKernelGGL(k_bls, dim3(1), dim3(DLS), 0, c.s, d.Rq, ld(), d.Gd, d.hv, d.hb, d.Tb, s, jj,
                           d.st->res);
        MARIE_HIP(hipGetLastError());
        return OK;
    }
    int enqueue_status() {
        MARIE_HIP(hipMemcpyAsync(d.hs, d.st, sizeof(DevStatus), hipMemcpyDeviceToHost, c.s));
        MARIE_HIP(hipEventRecord(d.ev, c.s));
        return OK;
    }
    int enqueue_step(int jj) {
        j = jj;
        phase = STEP;
        if (!use_graph) {
            const int st = step_body(jj);
            return st ? st : enqueue_status();
        }
        if ((int)d.graphs.size() <= jj) {
            d.graphs.resize(jj + 1, nullptr);
            d.gsrc.resize(jj + 1, nullptr);
        }
        if (!d.graphs[jj]) {
            MARIE_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
            const int st = step_body(jj);
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(c.s, &gr);
            if (st) {
                if (gr) (void)hipGraphDestroy(gr);
                return st;
            }
            if (e != hipSuccess) {
                set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                return ERR_HIP;
            }
            d.gsrc[jj] = gr;
            MARIE_HIP(hipGraphInstantiate(&d.graphs[jj], gr, nullptr, nullptr, 0));
        }
        MARIE_HIP(hipGraphLaunch(d.graphs[jj], c.s));
        return enqueue_status();
    }
    int start_cycle() {
        ++cycles;
        target.assign(s, 0.0);
        for (int q = 0; q < s; ++q) {
            double t = (mode == 1 ? 0.9 : 1.0) * tol * bn[q];
            if (mode == 1) t = std::max(t, inner_tol * rn[q]);
            target[q] = std::min(t, 0.5 * rn[q]);
        }
        int st = cholqr_dev(c.ws.R, c.ws.V, &d.st->brk0);
        if (st) return st;
        hipLaunchKernelGGL(k_ginit, dim3((unsigned)std::min<int64_t>(BG, (ld() * s + BT - 1) / BT)), dim3(BT), 0, c.s,
                           d.Rf, s, d.Gd, ld());
        MARIE_HIP(hipGetLastError());
        return enqueue_step(0);
    }
    int enqueue_cycle_end(int J) {
        steps = J;
        phase = CYCLE_END;
        const int K = J * s;
        const int64_t ns = c.n * s;
        const int g = grid_of(ns);
        hipLaunchKernelGGL(k_bsolve, dim3(s), dim3(BT), (size_t)2 * K * sizeof(double2), c.s, d.Rq, ld(), d.Gd, K,
                           c.ws.Sd);
        int st = vc(c, c.ws.D, c.ws.V, K, c.ws.Sd, K, s, 1.0, 0.0);
        if (st) return st;
        hipLaunchKernelGGL(k_add, dim3(g), dim3(BT), 0, c.s, X, c.ws.D, ns);
        st = apply(h, OP_AN, 0, PREC_F64, X, c.ws.T, s, c.s);
        if (st) return st;
        hipLaunchKernelGGL(k_sub, dim3(g), dim3(BT), 0, c.s, c.ws.R, B, c.ws.T, ns);
        hipLaunchKernelGGL(k_colnorm2, dim3(c.G, s), dim3(BT), 0, c.s, c.ws.R, c.n, c.ws.part, c.G);
        hipLaunchKernelGGL(k_colsum, dim3(s), dim3(BT), 0, c.s, c.ws.part, c.G, d.st->rn2);
        MARIE_HIP(hipGetLastError());
        return enqueue_status();
    }
    // set-up: the initial residual (synchronous), then the first cycle
    int init() {
        started = true;
        c.h = h;
        c.s = h->stream;
        c.n = h->nidx;
        c.ns = s;
        c.G = grid_of(c.n);
        f32 = mode == 1;
        if (!c.s) use_graph = false;
        int st = alloc();
        if (!st) st = ensure_work(h, s);
        if (!st) st = colnorms(c, B, bn);
        if (st) return fail(st);
        for (int q = 0; q < s; ++q)
            if (bn[q] == 0.0) {
                set_error("a right-hand side column is zero");
                return fail(ERR_ARG);
            }
        const int64_t ns = c.n * s;
        st = apply(h, OP_AN, 0, PREC_F64, X, c.ws.T, s, c.s);
        if (st) return fail(st);
        hipLaunchKernelGGL(k_sub, dim3(grid_of(ns)), dim3(BT), 0, c.s, c.ws.R, B, c.ws.T, ns);
        if (hipGetLastError() != hipSuccess) {
            set_error("kernel launch failed");
            return fail(ERR_HIP);
        }
        st = colnorms(c, c.ws.R, rn);
        if (st) return fail(st);
        hist.assign(1, worst_rel());
        if (worst_rel() <= tol || maxit < 1) {
            phase = DONE;
            return OK;
        }
        st = start_cycle();
        return st ? fail(st) : OK;
    }
    // the event of the enqueued work has completed
    int advance() {
        int st = OK;
        const DevStatus& hs = *d.hs;
        if (phase == STEP) {
            if (hs.brk0) {
                set_error("the block of residuals is rank deficient (dependent right-hand sides)");
                return fail(ERR_ARG);
            }
            double worst = 0.0;
            bool done = true;
            for (int q = 0; q < s; ++q) {
                worst = std::max(worst, hs.res[q] / bn[q]);
                if (hs.res[q] > target[q]) done = false;
            }
            hist.push_back(worst);
            const int J = j + 1;
            if (done || hs.brk || J == m) st = enqueue_cycle_end(J);
            else st = enqueue_step(J);
        } else if (phase == CYCLE_END) {
            for (int q = 0; q < s; ++q) rn[q] = std::sqrt(hs.rn2[q]);
            hist.push_back(worst_rel());
            if (worst_rel() <= tol || cycles >= maxit) phase = DONE;
            else st = start_cycle();
        }
        return st ? fail(st) : OK;
    }
    void results(int* flags, double* relres, int* iter) const {
        for (int q = 0; q < s; ++q) {
            relres[q] = rn[q] / bn[q];
            flags[q] = relres[q] <= tol ? 0 : 1;
        }
        iter[0] = steps;
        iter[1] = cycles;
    }
};

// Drive the solves to the end from one host thread: whichever solve's event
// has completed advances and enqueues its next piece of work on its stream.
void drive(std::vector<DevSolve*>& ps) {
    // each problem's work is enqueued with its handle's device current
    for (DevSolve* p : ps)
        if (p->phase != DevSolve::DONE) {
            (void)hipSetDevice(p->h->device);
            p->init();
        }
    for (;;) {
        int active = 0;
        bool moved = false;
        for (DevSolve* p : ps) {
            if (p->phase == DevSolve::DONE) continue;
            ++active;
            const hipError_t q = ps.size() == 1 ? hipEventSynchronize(p->d.ev) : hipEventQuery(p->d.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                set_error(std::string("block step: ") + hipGetErrorString(q));
                p->fail(ERR_HIP);
                continue;
            }
            (void)hipSetDevice(p->h->device);
            p->advance();
            moved = true;
        }
        if (!active) break;
        if (!moved) std::this_thread::yield();
    }
}

void set_dev_solve(DevSolve& p, Op* h, const void* B, void* X, int nrhs, int restart, double tol, int maxit,
                   int mode, double inner_tol, int options) {
    p.h = h;
    p.B = reinterpret_cast<const c2<double>*>(B);
    p.X = reinterpret_cast<c2<double>*>(X);
    p.s = nrhs;
    p.m = restart;
    p.mode = mode;
    p.maxit = maxit;
    p.tol = tol;
    p.inner_tol = inner_tol;
    p.ncgs = (options & MARIE_BLOCK_CGS2) ? 2 : 1;
    p.use_graph = (options & MARIE_BLOCK_GRAPH) != 0;
}

}  // namespace

int block_gmres(Op* h, const c2<double>* B, c2<double>* X, int s, int m, double tol, int maxit, int mode,
                double inner_tol, int options, int* flags, double* relres, int* iter, std::vector<double>& hist) {
    bool use_graph = (options & MARIE_BLOCK_GRAPH) != 0;
    const int ncgs = (options & MARIE_BLOCK_CGS2) ? 2 : 1;
    BCtx c;
    c.h = h;
    c.s = h->stream;
    c.n = h->nidx;
    c.ns = s;
    c.G = grid_of(c.n);
    if (!c.s) use_graph = false;   // the legacy default stream cannot be captured
    int st = alloc_block(c, m, mode);
    auto done = [&](int code) {
        (void)hipStreamSynchronize(c.s);
        c.ws.free_all();
        return code;
    };
    if (st) return done(st);
    st = ensure_work(h, s);
    if (st) return done(st);
    const int64_t ns = c.n * s;
    const int g = grid_of(ns);
    std::vector<double> bn, rn;
    st = colnorms(c, B, bn);
    if (st) return done(st);
    for (int q = 0; q < s; ++q)
        if (bn[q] == 0.0) {
            set_error("a right-hand side column is zero");
            return done(ERR_ARG);
        }
    // R = B - A X
    st = apply(h, OP_AN, 0, PREC_F64, X, c.ws.T, s, c.s);
    if (st) return done(st);
    hipLaunchKernelGGL(k_sub, dim3(g), dim3(BT), 0, c.s, c.ws.R, B, c.ws.T, ns);
    MARIE_HIP(hipGetLastError());
    st = colnorms(c, c.ws.R, rn);
    if (st) return done(st);
    auto worst_rel = [&]() {
        double w = 0.0;
        for (int q = 0; q < s; ++q) w = std::max(w, rn[q] / bn[q]);
        return w;
    };
    hist.assign(1, worst_rel());
    int cycles = 0, steps = 0;
    const char* ve = std::getenv("MARIE_VERBOSE");
    const bool verbose = ve && *ve && *ve != '0';
    while (worst_rel() > tol && cycles < maxit) {
        ++cycles;
        std::vector<double> target(s);
        for (int q = 0; q < s; ++q) {
            // stop the cycle at the global target (10 % margin in mode 1), at most
            // inner_tol relative to the cycle's residual in mode 1
            double t = (mode == 1 ? 0.9 : 1.0) * tol * bn[q];
            if (mode == 1) t = std::max(t, inner_tol * rn[q]);
            target[q] = std::min(t, 0.5 * rn[q]);
        }
        st = block_cycle(c, m, target, mode == 1, use_graph, ncgs, &steps, hist, bn);
        if (st) return done(st);
        hipLaunchKernelGGL(k_add, dim3(g), dim3(BT), 0, c.s, X, c.ws.D, ns);
        st = apply(h, OP_AN, 0, PREC_F64, X, c.ws.T, s, c.s);
        if (st) return done(st);
        hipLaunchKernelGGL(k_sub, dim3(g), dim3(BT), 0, c.s, c.ws.R, B, c.ws.T, ns);
        MARIE_HIP(hipGetLastError());
        st = colnorms(c, c.ws.R, rn);
        if (st) return done(st);
        hist.push_back(worst_rel());
        if (verbose) {
            std::fprintf(stderr, "[marie block gmres] cycle %d: %d block steps, worst relres %.3e\n", cycles, steps,
                         worst_rel());
            std::fflush(stderr);
        }
    }
    for (int q = 0; q < s; ++q) {
        relres[q] = rn[q] / bn[q];
        flags[q] = relres[q] <= tol ? 0 : 1;
    }
    iter[0] = steps;
    iter[1] = cycles;
    return done(OK);
}

}  // namespace marie

using namespace marie;

extern "C" int marie_vie_block_gmres_device(marie_vie_op* hh, const void* B, void* X, int nrhs, int restart,
                                            double tol, int maxit, int mode, double inner_tol, int options,
                                            int* flag, double* relres, int* iter, double* resvec, int resvec_cap,
                                            int* nres) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (options & MARIE_BLOCK_COLUMNS) {
        // the columns one at a time, each a marie_vie_gmres_device solve in
        // `mode` (restart = its cycle length); iter = {the largest last-cycle
        // count, the cycles of all columns}; resvec: each column's final relres
        if (!B || !X || !flag || !relres || !iter || nrhs < 1) {
            set_error("invalid block gmres arguments");
            return MARIE_ERR_ARG;
        }
        const int64_t n = h->nidx;
        iter[0] = iter[1] = 0;
        for (int q = 0; q < nrhs; ++q) {
            int it[2] = {0, 0};
            const int st = marie_vie_gmres_device(hh, (const c2<double>*)B + q * n, (c2<double>*)X + q * n, restart,
                                                  tol, maxit, mode, inner_tol, flag + q, relres + q, it, nullptr, 0,
                                                  nullptr);
            if (st) return st;
            iter[0] = std::max(iter[0], it[0]);
            iter[1] += it[1];
            if (resvec && q < resvec_cap) resvec[q] = relres[q];
        }
        if (nres) *nres = std::min(nrhs, resvec ? resvec_cap : 0);
        return MARIE_OK;
    }
    if (!B || !X || !flag || !relres || !iter || nrhs < 1 || nrhs > 64 || restart < 1 ||
        (int64_t)(restart + 1) * nrhs > 2048 || maxit < 1 || !(tol > 0) || (mode != 0 && mode != 1)) {
        set_error("invalid block gmres arguments (nrhs 1..64, (restart+1)*nrhs <= 2048, mode 0 or 1)");
        return MARIE_ERR_ARG;
    }
    if ((options & MARIE_BLOCK_DEVICE) && nrhs > DS) {
        set_error("device-driven block steps take at most 32 right-hand sides");
        return MARIE_ERR_ARG;
    }
    if (!(h->precisions & PREC_F64) || (mode == 1 && !(h->precisions & PREC_F32))) {
        set_error("block gmres needs an FP64 operator (and FP32 for mode 1)");
        return MARIE_ERR_STATE;
    }
    if (!h->have_fN || !h->have_mat || !h->have_idx) {
        set_error("operator, index or material not set");
        return MARIE_ERR_STATE;
    }
    if (h->nidx == 0) {
        for (int q = 0; q < nrhs; ++q) { flag[q] = 0; relres[q] = 0.0; }
        iter[0] = iter[1] = 0;
        if (nres) *nres = 0;
        return MARIE_OK;
    }
    std::vector<double> hist;
    if (options & MARIE_BLOCK_DEVICE) {
        DevSolve p;
        set_dev_solve(p, h, B, X, nrhs, restart, tol, maxit, mode, inner_tol, options);
        std::vector<DevSolve*> ps(1, &p);
        drive(ps);
        if (p.err) {
            set_error(p.msg);
            return p.err;
        }
        p.results(flag, relres, iter);
        hist = p.hist;
    } else {
        const int st = block_gmres(h, (const c2<double>*)B, (c2<double>*)X, nrhs, restart, tol, maxit, mode,
                                   inner_tol, options, flag, relres, iter, hist);
        if (st) return st;
    }
    if (resvec && resvec_cap > 0)
        for (int i = 0; i < (int)hist.size() && i < resvec_cap; ++i) resvec[i] = hist[i];
    if (nres) *nres = (int)hist.size();
    return MARIE_OK;
}

extern "C" int marie_vie_block_gmres(marie_vie_op* hh, const double* B, double* X, int nrhs, int restart, double tol,
                                     int maxit, int mode, double inner_tol, int options, int* flag, double* relres,
                                     int* iter, double* resvec, int resvec_cap, int* nres) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    if (!B || !X || nrhs < 1) { set_error("null block or nrhs < 1"); return MARIE_ERR_ARG; }
    const size_t n = (size_t)h->nidx * nrhs;
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(B[i]) || !std::isfinite(X[i])) {
            set_error("right-hand sides or initial guesses have non-finite entries");
            return MARIE_ERR_NONFINITE;
        }
    void *db = nullptr, *dx = nullptr;
    if (n) {
        if (hipMalloc(&db, n * sizeof(c2<double>)) != hipSuccess || hipMalloc(&dx, n * sizeof(c2<double>)) != hipSuccess) {
            (void)hipGetLastError();
            if (db) (void)hipFree(db);
            set_error("out of device memory");
            return MARIE_ERR_ALLOC;
        }
        MARIE_HIP(hipMemcpy(db, B, n * sizeof(c2<double>), hipMemcpyHostToDevice));
        MARIE_HIP(hipMemcpy(dx, X, n * sizeof(c2<double>), hipMemcpyHostToDevice));
    }
    int st = marie_vie_block_gmres_device(hh, db, dx, nrhs, restart, tol, maxit, mode, inner_tol, options, flag,
                                          relres, iter, resvec, resvec_cap, nres);
    if (st == MARIE_OK && n && hipMemcpy(X, dx, n * sizeof(c2<double>), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("D2H copy failed");
        st = MARIE_ERR_HIP;
    }
    if (db) (void)hipFree(db);
    if (dx) (void)hipFree(dx);
    return st;
}

extern "C" int marie_vie_block_gmres_batch(marie_vie_op* const* ops, int nitems, const void* const* B, void* const* X,
                                           const int* nrhs, int restart, double tol, int maxit, int mode,
                                           double inner_tol, int options, int* flag, double* relres, int* iter,
                                           int* status) {
    if (!ops || !B || !X || !nrhs || !flag || !relres || !iter || !status || nitems < 1 || nitems > 256) {
        set_error("invalid batch arguments (1..256 problems)");
        return MARIE_ERR_ARG;
    }
    std::vector<int64_t> off(nitems + 1, 0);
    for (int i = 0; i < nitems; ++i) {
        if (!ops[i] || nrhs[i] < 1) {
            set_error("batch problem " + std::to_string(i) + ": null handle or no right-hand side");
            return MARIE_ERR_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (ops[j] == ops[i]) {
                set_error("batch problems " + std::to_string(j) + " and " + std::to_string(i) +
                          " share a handle (each needs its own stream and work buffers)");
                return MARIE_ERR_ARG;
            }
        off[i + 1] = off[i] + nrhs[i];
    }
    if (options & MARIE_BLOCK_DEVICE) {
        // one host thread, every problem's steps on its own stream
        for (int i = 0; i < nitems; ++i) {
            const Op* hi = reinterpret_cast<const Op*>(ops[i]);
            if (!B[i] || !X[i] || nrhs[i] > DS || restart < 1 || (int64_t)(restart + 1) * nrhs[i] > 2048 ||
                maxit < 1 || !(tol > 0) || (mode != 0 && mode != 1)) {
                set_error("batch problem " + std::to_string(i) +
                          ": invalid arguments (nrhs <= 32 with device-driven steps, (restart+1)*nrhs <= 2048, "
                          "mode 0 or 1)");
                return MARIE_ERR_ARG;
            }
            if (!(hi->precisions & PREC_F64) || (mode == 1 && !(hi->precisions & PREC_F32)) || !hi->have_fN ||
                !hi->have_mat || !hi->have_idx) {
                set_error("batch problem " + std::to_string(i) +
                          ": operator, index or material not set, or a precision missing");
                return MARIE_ERR_STATE;
            }
        }
        std::vector<DevSolve> ps(nitems);
        std::vector<DevSolve*> pp(nitems);
        for (int i = 0; i < nitems; ++i) {
            Op* hi = reinterpret_cast<Op*>(ops[i]);
            MARIE_HIP(hipSetDevice(hi->device));
            set_dev_solve(ps[i], hi, B[i], X[i], nrhs[i], restart, tol, maxit, mode, inner_tol, options);
            if (hi->nidx == 0) {
                // no DOFs: converged with zero iterations, as the single-problem entry
                ps[i].phase = DevSolve::DONE;
                ps[i].bn.assign(nrhs[i], 1.0);
                ps[i].rn.assign(nrhs[i], 0.0);
            }
            pp[i] = &ps[i];
        }
        drive(pp);
        int first = MARIE_OK;
        std::string fmsg;
        for (int i = 0; i < nitems; ++i) {
            status[i] = ps[i].err;
            if (ps[i].err) {
                if (first == MARIE_OK) {
                    first = ps[i].err;
                    fmsg = "batch problem " + std::to_string(i) + ": " + ps[i].msg;
                }
                continue;
            }
            ps[i].results(flag + off[i], relres + off[i], iter + 2 * i);
        }
        if (first != MARIE_OK) set_error(fmsg);
        return first;
    }
    // largest problems first, so the small ones fill in around them
    std::vector<int> order(nitems);
    for (int i = 0; i < nitems; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const Op* ha = reinterpret_cast<const Op*>(ops[a]);
        const Op* hb = reinterpret_cast<const Op*>(ops[b]);
        return ha->nidx * nrhs[a] > hb->nidx * nrhs[b];
    });
    std::vector<std::string> msg(nitems);
    std::atomic<int> next(0);
    auto worker = [&]() {
        for (int t = next++; t < nitems; t = next++) {
            const int i = order[t];
            status[i] = marie_vie_block_gmres_device(ops[i], B[i], X[i], nrhs[i], restart, tol, maxit, mode, inner_tol,
                                                     options, flag + off[i], relres + off[i], iter + 2 * i, nullptr,
                                                     0, nullptr);
            if (status[i] != MARIE_OK) msg[i] = get_error();
        }
    };
    const int nth = nitems < 8 ? nitems : 8;
    std::vector<std::thread> th;
    for (int k = 1; k < nth; ++k) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    for (int i = 0; i < nitems; ++i)
        if (status[i] != MARIE_OK) {
            set_error("batch problem " + std::to_string(i) + ": " + msg[i]);
            return status[i];
        }
    return MARIE_OK;
}
