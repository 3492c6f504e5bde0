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
                                            c2<double>* __restrict__ part, int G) {
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
                                                   c2<double>* __restrict__ C, int64_t ldc) {
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
                                           int64_t n, double alpha, double beta) {
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

// C (P x Q, ld ldc) = V^H W over n rows
int vhw(BCtx& c, const c2<double>* V, int P, const c2<double>* W, int Q, c2<double>* C, int64_t ldc) {
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
        hipLaunchKernelGGL(k_vhw<4>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G);
    else if (wpt == 2)
        hipLaunchKernelGGL(k_vhw<2>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G);
    else
        hipLaunchKernelGGL(k_vhw<1>, grid, dim3(BT), 0, c.s, V, c.n, P, W, c.n, Q, c.n, c.ws.vpart, G);
    hipLaunchKernelGGL(k_vhw_reduce, dim3((P * Q + BT / 64 - 1) / (BT / 64)), dim3(BT), 0, c.s, c.ws.vpart, P, Q, G,
                       C, ldc);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// out (n x Q) = beta * out + alpha * V (n x P) C (P x Q, ld ldc)
int vc(BCtx& c, c2<double>* out, const c2<double>* V, int P, const c2<double>* C, int64_t ldc, int Q, double alpha,
       double beta) {
    const dim3 g((unsigned)((c.n + BT - 1) / BT), (unsigned)((Q + CQ - 1) / CQ));
    hipLaunchKernelGGL(k_vc, g, dim3(BT), 0, c.s, out, c.n, V, c.n, P, C, ldc, Q, c.n, alpha, beta);
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
    const int st = block_gmres(h, (const c2<double>*)B, (c2<double>*)X, nrhs, restart, tol, maxit, mode, inner_tol,
                               options, flag, relres, iter, hist);
    if (st) return st;
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
