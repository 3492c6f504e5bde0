// Block-orthogonalised GMRES cycles on complex64 products inside FP64
// iterative refinement: marie_vie_gmres_device mode 3.
//
// The refinement is that of modes 1-2 (pgmres.m's flag/relres on the FP64
// true residual, the FP64 residual replaced at every restart); what differs
// is the inner cycle. Classical Gram-Schmidt (pgmres_iter.m:106-107) reads
// the whole basis twice per product; here the basis grows by blocks of up
// to SM = 8 columns:
//
//   b_0 = q_last; for j < S: w = A b_j, b_{j+1} = w minus its projection on
//        b_{j-1}, b_j, normalised              banded local Arnoldi (the block only)
//   P = [b_1 .. b_S]
//   [C; G0] = [Q P]^H P                        one pass over Q and P (matrix cores)
//   R1 = chol(G0 - C^H C)                      on the device (block Gram-Schmidt with
//   P  = (P - Q C) R1^{-1}                     the Pythagorean Gram, one pass)
//   R2 = chol(P^H P),  P = P R2^{-1}           second CholeskyQR pass, skipped on the
//                                              device when every column kept 1e-2 of its
//                                              squared norm in the first (then R2 = I)
//
// so the whole basis is read twice per block instead of twice per product,
// while every product still takes a unit vector orthogonal to the two
// columns before it (a well-conditioned block, unlike a power basis). With
// the local relation A b_j = [b_0 .. b_{j+1}] h_j (h_j nonzero in rows j-1 ..
// j+1 only) and P = Q C + Q_new R2 R1, the Arnoldi
// relation A Q_k = Q_{k+1} H follows column by column (host, FP64); the
// cycle's least squares min || beta e_1 - H y || is updated with Givens
// rotations on the host. A CholeskyQR pivot below 1e-10 of the column's
// squared norm ends the block at that column; the next block continues from
// the last accepted one.
//
// Every dot product accumulates in FP64 (the complex64 operands convert
// exactly); the basis is stored in complex64.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"
#include "vie_reduce.hpp"
#include "vie_dense.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <string>
#include <vector>

namespace marie {

namespace {

using cd = std::complex<double>;
constexpr int ST = 256;          // threads per block
constexpr int SG = 1024;         // max blocks of a reduction
constexpr int SM = 8;            // largest block of products (columns of P)
constexpr int SKR = 64;          // rows per staged chunk of the dot kernel
constexpr int SGD = 1024;        // most row blocks of the dot kernel
constexpr int kMaxBasis = 512;   // most basis columns a combination reads
constexpr int kSchRows = 64;     // basis rows of C per LDS stage of the Cholesky kernel

__device__ __forceinline__ double2 ld2(const c2<float>* p) {
    const c2<float> v = *p;
    return double2{(double)v.x, (double)v.y};
}
__device__ __forceinline__ double2 ld2(const c2<double>* p) {
    const c2<double> v = *p;
    return double2{v.x, v.y};
}
// one complex64 basis element, non-temporal: the block passes stream the
// whole basis (far larger than the last-level cache) once each, and the
// operator's octant and work buffers keep their place in the cache
__device__ __forceinline__ c2<float> ldf_nt(const c2<float>* p) {
    typedef float v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return c2<float>{v.x, v.y};
}

// part[(p * SM + q) * G + rb] = sum over the row chunks rb, rb + G, ... of
// conj(V_p) W_q, p < P, q < Qn (<= SM). V, W complex64 columns (stride
// ldq). A workgroup (4 waves) takes 64 columns of V and the W block; each
// chunk of SKR rows is staged in LDS, then the 4-row k-steps run on the
// matrix cores (v_mfma_f64_16x16x4f64) with each tile of 8 W columns packed
// as the 16 real columns [Re W | Im W]: D1 = Re(V)^T [Re W | Im W], D2 =
// Im(V)^T [Re W | Im W], so conj(V)^T W = (D1[:, q] + D2[:, 8+q]) + j (D1[:,
// 8+q] - D2[:, q]) takes two real products per k-step and tile. When the tile holds fewer than 64
// columns, the waves of a 16-column subtile split its k-steps and add their
// sums in LDS. Operand layout: lane i + 16 k holds A[i][k] and B[k][i];
// D[r][c] is element r / 4 of lane c + 16 (r % 4).
typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int STP = 64;

__global__ __launch_bounds__(ST) void k_sdot(const c2<float>* __restrict__ V, int P, const c2<float>* __restrict__ W,
                                             int Qn, int64_t ldq, int64_t n, c2<double>* __restrict__ part, int G,
                                             const int* __restrict__ skip) {
    if (skip && *skip) return;
    static_assert(SM % 8 == 0, "the packed W operand holds 8 complex columns per tile");
    constexpr int NWT = SM / 8;                      // W tiles of 8 complex columns
    __shared__ __attribute__((aligned(32))) c2<float> Vs[SKR][STP + 1];
    __shared__ c2<float> Ws[SKR][SM + 1];
    const int p0 = blockIdx.x * STP;
    const int rb = blockIdx.y;
    const int pv = P - p0 < STP ? P - p0 : STP;
    const int nsub = (pv + 15) / 16;                 // 16-column subtiles with data
    const int wps = nsub == 1 ? 4 : (nsub == 2 ? 2 : 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = wave / wps, kpart = wave % wps;
    const bool active = sub < nsub;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t nchunk = (n + SKR - 1) / SKR;
    constexpr int LV = SKR * STP / ST;               // V elements staged per thread and chunk
    constexpr int LW = SKR * SM / ST;                // W elements staged per thread and chunk
    c2<float> vreg[LV], wreg[LW];
    auto fetch = [&](int64_t ch) {
        const int64_t r0 = ch * SKR;
#pragma unroll
        for (int k = 0; k < LV; ++k) {
            const int i = threadIdx.x + ST * k;
            const int r = i % SKR, pc = i / SKR;
            const int64_t e = r0 + r;
            vreg[k] = (pc < pv && e < n) ? ldf_nt(V + (int64_t)(p0 + pc) * ldq + e) : c2<float>{0.f, 0.f};
        }
#pragma unroll
        for (int k = 0; k < LW; ++k) {
            const int i = threadIdx.x + ST * k;
            const int r = i % SKR, qc = i / SKR;
            const int64_t e = r0 + r;
            wreg[k] = (qc < Qn && e < n) ? W[(int64_t)qc * ldq + e] : c2<float>{0.f, 0.f};
        }
    };
    dx4 d1[NWT], d2[NWT];
#pragma unroll
    for (int t = 0; t < NWT; ++t) {
        d1[t] = dx4{0, 0, 0, 0};
        d2[t] = dx4{0, 0, 0, 0};
    }
    if (rb < nchunk) fetch(rb);
    for (int64_t ch = rb; ch < nchunk; ch += G) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < LV; ++k) {
            const int i = threadIdx.x + ST * k;
            Vs[i % SKR][i / SKR] = vreg[k];
        }
#pragma unroll
        for (int k = 0; k < LW; ++k) {
            const int i = threadIdx.x + ST * k;
            Ws[i % SKR][i / SKR] = wreg[k];
        }
        __syncthreads();
        if (ch + G < nchunk) fetch(ch + G);
        if (active) {
            for (int ks = kpart; ks < SKR / 4; ks += wps) {
                const c2<float> a = Vs[4 * ks + lk][16 * sub + li];
#pragma unroll
                for (int t = 0; t < NWT; ++t) {
                    const c2<float> w = Ws[4 * ks + lk][8 * t + (li & 7)];
                    const double bv = li < 8 ? (double)w.x : (double)w.y;
                    d1[t] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a.x, bv, d1[t], 0, 0, 0);
                    d2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a.y, bv, d2[t], 0, 0, 0);
                }
            }
        }
    }
    if (wps > 1) {
        // the k-step sums of a subtile's waves, added in a fixed order (the
        // staging area is free after the barrier)
        static_assert(4 * 64 * 2 * NWT * sizeof(dx4) <= sizeof(Vs), "reduction area");
        __syncthreads();
        dx4* red = reinterpret_cast<dx4*>(&Vs[0][0]);     // 4 waves x 64 lanes x 2 NWT dx4
#pragma unroll
        for (int t = 0; t < NWT; ++t) {
            red[((wave * 64 + lane) * NWT + t) * 2] = d1[t];
            red[((wave * 64 + lane) * NWT + t) * 2 + 1] = d2[t];
        }
        __syncthreads();
        if (kpart != 0) return;
        for (int w2 = 1; w2 < wps; ++w2)
#pragma unroll
            for (int t = 0; t < NWT; ++t) {
                d1[t] += red[(((wave + w2) * 64 + lane) * NWT + t) * 2];
                d2[t] += red[(((wave + w2) * 64 + lane) * NWT + t) * 2 + 1];
            }
    }
    if (!active) return;
#pragma unroll
    for (int t = 0; t < NWT; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // lanes li < 8 combine with lane li + 8 of the same k group
            const double d1h = __shfl_down(d1[t][e], 8, 16), d2h = __shfl_down(d2[t][e], 8, 16);
            const int pr = 16 * sub + 4 * e + lk, q = 8 * t + li;
            if (li < 8 && pr < pv && q < Qn)
                part[((int64_t)(p0 + pr) * SM + q) * G + rb] = c2<double>{d1[t][e] + d2h, d1h - d2[t][e]};
        }
}

// C[p * SM + q] = sum_rb part[(p * SM + q) G + rb] (q < Qn; columns past Qn
// are set to zero): one wave per output, fixed order
__global__ __launch_bounds__(ST) void k_sdot_reduce(const c2<double>* __restrict__ part, int P, int Qn, int G,
                                                    c2<double>* __restrict__ C, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int o = blockIdx.x * (ST / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= P * SM) return;
    const int q = o % SM;
    double x = 0.0, y = 0.0;
    if (q < Qn) {
        const c2<double>* src = part + (int64_t)o * G;
        for (int rb = lane; rb < G; rb += 64) {
            x += src[rb].x;
            y += src[rb].y;
        }
        for (int s = 32; s > 0; s >>= 1) {
            x += __shfl_down(x, s, 64);
            y += __shfl_down(y, s, 64);
        }
    }
    if (lane == 0) C[o] = c2<double>{x, y};
}

// CholeskyQR factor of the block's Gram: G[c][d] = C[(k + c) SM + d] minus
// (proj) sum_{i<k} conj(C[i SM + c]) C[i SM + d] (the Pythagorean form of
// (P - Q C)^H (P - Q C) with C = Q^H P). R (upper, column-major SM x SM) with
// G = R^H R, Rinv = R^{-1}; the factor stops at the first column whose pivot
// is not above thresh * C[(k + j) SM + j] (the column's squared norm before
// the projection): nok[0] = accepted columns; R and Rinv are zero past them.
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 zcmul(double2 a, double2 b) {      // conj(a) b
    return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// squared-norm fraction every column of a block must keep after its
// projection for the second CholeskyQR pass to be skipped
constexpr double kWell = 1e-2;

__global__ __launch_bounds__(ST) void k_schol(const c2<double>* __restrict__ C, int k, int sb, int proj, double thresh,
                                              c2<double>* __restrict__ R, c2<double>* __restrict__ Rinv,
                                              int* __restrict__ nok, const int* __restrict__ skip,
                                              int* __restrict__ well) {
    static_assert(SM * SM <= ST, "one thread per entry of the block's Gram");
    const int t = threadIdx.x;
    if (skip && *skip) {
        // the second factor of a block whose first one kept its columns well
        // apart from the basis: the identity
        if (t < SM * SM) {
            const c2<double> v = {t % SM == t / SM ? 1.0 : 0.0, 0.0};
            R[t] = v;
            Rinv[t] = v;
        }
        if (t == 0) nok[0] = sb;
        return;
    }
    __shared__ double2 Gs[SM][SM];
    __shared__ double2 Rs[SM][SM];
    __shared__ double2 Is[SM][SM];
    __shared__ double d0[SM];
    __shared__ double rinv[SM];
    __shared__ double2 Cs[kSchRows][SM];
    __shared__ int ok;
    // entry (c, d) of C^H C over the basis rows i < k: the rows staged through
    // LDS kSchRows at a time (coalesced loads), one thread per entry
    const int c = t / SM, d = t % SM;
    double2 acc = {0.0, 0.0};
    if (proj) {
        for (int i0 = 0; i0 < k; i0 += kSchRows) {
            const int nr = min(kSchRows, k - i0);
            for (int e = t; e < nr * SM; e += ST) {
                const c2<double> v = C[(int64_t)i0 * SM + e];
                Cs[e / SM][e % SM] = double2{v.x, v.y};
            }
            __syncthreads();
            if (t < SM * SM && c < sb && d < sb) {
                for (int i = 0; i < nr; ++i) {
                    const double2 p = zcmul(Cs[i][c], Cs[i][d]);
                    acc.x += p.x;
                    acc.y += p.y;
                }
            }
            __syncthreads();
        }
    }
    if (t < SM * SM) {
        double2 g = {0.0, 0.0};
        if (c < sb && d < sb) {
            const c2<double> v = C[(int64_t)(k + c) * SM + d];
            g = double2{v.x - acc.x, v.y - acc.y};
            if (c == d) d0[c] = v.x;
        }
        Gs[c][d] = g;
        Rs[c][d] = double2{0.0, 0.0};
        Is[c][d] = double2{0.0, 0.0};
    }
    if (t == 0) ok = 0;
    __syncthreads();
    // right-looking Cholesky G = R^H R (R's diagonal real and positive); the
    // factor stops at the first pivot not above thresh * the column's squared
    // norm before the projection
    for (int j = 0; j < sb; ++j) {
        if (t == 0) {
            const double piv = Gs[j][j].x;
            if (piv > thresh * d0[j]) {
                const double rjj = sqrt(piv);
                Rs[j][j] = double2{rjj, 0.0};
                rinv[j] = 1.0 / rjj;
                ok = j + 1;
            }
        }
        __syncthreads();
        if (ok != j + 1) break;
        if (t > j && t < sb) Rs[j][t] = double2{Gs[j][t].x * rinv[j], Gs[j][t].y * rinv[j]};
        __syncthreads();
        if (t < SM * SM) {
            const int a = t / SM, b = t % SM;
            if (a > j && b >= a && b < sb) {
                const double2 p = zcmul(Rs[j][a], Rs[j][b]);
                Gs[a][b].x -= p.x;
                Gs[a][b].y -= p.y;
            }
        }
        __syncthreads();
    }
    const int nk = ok;
    // column l of R^{-1} (leading nk x nk upper triangle) by back substitution
    if (t < nk) {
        const int l = t;
        Is[l][l] = double2{rinv[l], 0.0};
        for (int i = l - 1; i >= 0; --i) {
            double2 v = {0.0, 0.0};
            for (int jj = i + 1; jj <= l; ++jj) {
                const double2 p = zmul(Rs[i][jj], Is[jj][l]);
                v.x += p.x;
                v.y += p.y;
            }
            Is[i][l] = double2{-v.x * rinv[i], -v.y * rinv[i]};
        }
    }
    __syncthreads();
    if (t < SM * SM) {
        const int c = t % SM, d = t / SM;     // column-major: entry (c, d) at d * SM + c
        const bool in = c < nk && d < nk;
        R[d * SM + c] = in ? c2<double>{Rs[c][d].x, Rs[c][d].y} : c2<double>{0.0, 0.0};
        Rinv[d * SM + c] = in ? c2<double>{Is[c][d].x, Is[c][d].y} : c2<double>{0.0, 0.0};
    }
    if (t == 0) {
        nok[0] = nk;
        if (well) {
            // every pivot kept at least kWell of its column's squared norm:
            // the Pythagorean Gram lost at most a factor 1/kWell to
            // cancellation and one CholeskyQR pass leaves the block orthonormal
            // to the basis' own accuracy
            double rmin = 1.0;
            for (int jj = 0; jj < nk; ++jj) rmin = fmin(rmin, Rs[jj][jj].x * Rs[jj][jj].x / d0[jj]);
            *well = (nk == sb && rmin >= kWell) ? 1 : 0;
        }
    }
}

// P = (P - Q C) Rinv on the block's columns P_c = Q[k + c], c < sb (one thread
// per row; Q columns i < k with coefficients C[i SM + c] read as uniform
// operands; Rinv upper triangular, column-major). FP64 arithmetic on the
// complex64 data, complex64 result in place.
__global__ __launch_bounds__(ST) void k_supdate(c2<float>* __restrict__ Q, int64_t ldq, int k, int sb,
                                                const c2<double>* __restrict__ C, const c2<double>* __restrict__ Rinv,
                                                int proj, int64_t n, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int64_t e = blockIdx.x * (int64_t)ST + threadIdx.x;
    if (e >= n) return;
    // the block's columns: all SM loads issued before the first is used
    // (columns past sb read column k and are zeroed)
    double ax[SM], ay[SM];
    c2<float> bq[SM];
#pragma unroll
    for (int c = 0; c < SM; ++c) bq[c] = Q[(int64_t)(k + (c < sb ? c : 0)) * ldq + e];
#pragma unroll
    for (int c = 0; c < SM; ++c) {
        ax[c] = c < sb ? (double)bq[c].x : 0.0;
        ay[c] = c < sb ? (double)bq[c].y : 0.0;
    }
    if (proj) {
        // basis rows QU at a time, their loads in flight together (one load
        // per wave in flight left the projection latency bound); the sums
        // run over i in the same order as one row at a time
        constexpr int QU = 4;
        auto project = [&](const c2<float> qf, int i) {
            const double2 q = double2{(double)qf.x, (double)qf.y};
            const c2<double>* ci = C + (int64_t)i * SM;
#pragma unroll
            for (int c = 0; c < SM; ++c) {
                const c2<double> cc = ci[c];
                ax[c] -= q.x * cc.x - q.y * cc.y;
                ay[c] -= q.x * cc.y + q.y * cc.x;
            }
        };
        int i = 0;
        for (; i + QU <= k; i += QU) {
            c2<float> qf[QU];
#pragma unroll
            for (int u = 0; u < QU; ++u) qf[u] = ldf_nt(Q + (int64_t)(i + u) * ldq + e);
#pragma unroll
            for (int u = 0; u < QU; ++u) project(qf[u], i + u);
        }
        for (; i < k; ++i) project(ldf_nt(Q + (int64_t)i * ldq + e), i);
    }
#pragma unroll
    for (int d = SM - 1; d >= 0; --d) {
        if (d < sb) {
            double ox = 0.0, oy = 0.0;
#pragma unroll
            for (int c = 0; c <= d; ++c) {
                const c2<double> r = Rinv[d * SM + c];
                ox += ax[c] * r.x - ay[c] * r.y;
                oy += ax[c] * r.y + ay[c] * r.x;
            }
            Q[(int64_t)(k + d) * ldq + e] = c2<float>{(float)ox, (float)oy};
        }
    }
}

// ---- a block's local Arnoldi: each product's output orthogonalised against
// the block's columns Q[c0 .. c0 + nl) (the column it multiplied and the one
// before it, the cycle's last multiplied column opening the block), so every
// product takes a unit vector; the block's columns meet each other and the
// rest of the basis in the block Gram-Schmidt pass.
constexpr int SL = SM + 1;       // columns b_0 .. b_SM of a block's local relation
// columns each product is orthogonalised against inside its block: the last
// two, b_{j-1} and b_j (the block's CholeskyQR makes the rest orthogonal;
// against all of b_0 .. b_j the 128^3 solve needed the same 166 products)
constexpr int kBand = 2;

__device__ __forceinline__ float4 ld4(const c2<float>* p) { return *reinterpret_cast<const float4*>(p); }
// The block's columns are read non-temporally as well: through the caches
// they pushed the operator's octant out between products (the complex64 z
// pass inside the solve 146 -> 138 us at 256^3, the local dots 21 -> 23 us).
__device__ __forceinline__ float4 ld4nt(const c2<float>* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
    return float4{v.x, v.y, v.z, v.w};
}

// part[i * G + blk] = block sums of conj(Q_{c0+i}) w over its rows, i < nl,
// and part[nl * G + blk] = block sum of |w|^2.
// Columns start on 16-byte boundaries (even ldq): each thread reads two
// adjacent elements per 16-byte load; the last element of an odd n is added
// by thread 0 of block 0.
__global__ __launch_bounds__(ST) void k_ldot(const c2<float>* __restrict__ Q, int64_t ldq, int c0, int nl,
                                             const c2<float>* __restrict__ w, int64_t n, c2<double>* __restrict__ part,
                                             int G) {
    static_assert(ST == 256, "the block sums read four LDS entries per lane");
    static_assert(kBand <= SL, "the band lies inside the block");
    __shared__ double sh[(2 * kBand + 1) * ST];
    double acc[2 * kBand + 1];
#pragma unroll
    for (int q = 0; q < 2 * kBand + 1; ++q) acc[q] = 0.0;
    const int64_t n2 = n >> 1;
    const c2<float>* qb = Q + (int64_t)c0 * ldq;
    for (int64_t e = blockIdx.x * (int64_t)ST + threadIdx.x; e < n2; e += (int64_t)G * ST) {
        // every column's load is issued before the first product uses one
        float4 a[kBand];
#pragma unroll
        for (int q = 0; q < kBand; ++q)
            if (q < nl) a[q] = ld4nt(qb + (int64_t)q * ldq + 2 * e);
        const float4 wv = ld4(w + 2 * e);
        acc[2 * kBand] += (double)wv.x * wv.x + (double)wv.y * wv.y + (double)wv.z * wv.z + (double)wv.w * wv.w;
#pragma unroll
        for (int q = 0; q < kBand; ++q) {
            if (q < nl) {
                acc[2 * q] += (double)a[q].x * wv.x + (double)a[q].y * wv.y + (double)a[q].z * wv.z +
                              (double)a[q].w * wv.w;
                acc[2 * q + 1] += (double)a[q].x * wv.y - (double)a[q].y * wv.x + (double)a[q].z * wv.w -
                                  (double)a[q].w * wv.z;
            }
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double2 wv = ld2(w + n - 1);
        acc[2 * kBand] += wv.x * wv.x + wv.y * wv.y;
        for (int q = 0; q < nl; ++q) {
            const double2 a = ld2(Q + (int64_t)(c0 + q) * ldq + n - 1);
            acc[2 * q] += a.x * wv.x + a.y * wv.y;
            acc[2 * q + 1] += a.x * wv.y - a.y * wv.x;
        }
    }
    // block sums through LDS: row q of sh holds every thread's acc[q]; wave w
    // adds rows w, w + 4, ... (four LDS reads per lane, then a shuffle tree)
    // and its lane 0 writes the row's real or imaginary part
#pragma unroll
    for (int q = 0; q < 2 * kBand + 1; ++q) sh[q * ST + threadIdx.x] = acc[q];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* pd = reinterpret_cast<double*>(part);
    for (int q = wave; q < 2 * kBand + 1; q += ST / 64) {
        const bool used = q == 2 * kBand || (q >> 1) < nl;
        if (!used) continue;
        double v = sh[q * ST + lane] + sh[q * ST + lane + 64] + sh[q * ST + lane + 128] + sh[q * ST + lane + 192];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) {
            const int64_t slot = q == 2 * kBand ? (int64_t)nl * G + blockIdx.x : (int64_t)(q >> 1) * G + blockIdx.x;
            pd[2 * slot + (q == 2 * kBand ? 0 : (q & 1))] = v;
            if (q == 2 * kBand) pd[2 * slot + 1] = 0.0;
        }
    }
}

// h[i] = sum_b part[i * G + b], fixed order; one block per i < nl, and block
// nl writes the sum of row nl (|w|^2) to *w2
__global__ __launch_bounds__(ST) void k_lsum(const c2<double>* __restrict__ part, int G, int nl,
                                             c2<double>* __restrict__ h, c2<double>* __restrict__ w2) {
    __shared__ double sh[ST / 64 * 2];
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < G; b += ST) {
        const c2<double> p = part[(int64_t)blockIdx.x * G + b];
        acc[0] += p.x;
        acc[1] += p.y;
    }
    block_reduce<2>(acc, sh);
    if (threadIdx.x == 0) (blockIdx.x < (unsigned)nl ? h[blockIdx.x] : *w2) = c2<double>{acc[0], acc[1]};
}

// out = (w - sum_{i<nl} Q_{c0+i} h[i]) / beta (FP64 arithmetic, complex64
// result), with beta^2 = |w|^2 - sum |h_i|^2 (the columns are orthonormal:
// the Pythagorean norm of the projected vector, so no second pass over it);
// block 0 stores beta in h[nl], the local Hessenberg column's subdiagonal.
// beta^2 below 1e-12 |w|^2 (w in the block's span to FP64 accuracy) gives a
// zero column, where the block's CholeskyQR then stops. Two elements per
// thread per 16-byte access (even ldq), the last element of an odd n by
// thread 0 of block 0.
__global__ __launch_bounds__(ST) void k_lcomb(c2<float>* __restrict__ out, const c2<float>* __restrict__ w,
                                              const c2<float>* __restrict__ Q, int64_t ldq, int c0, int nl,
                                              c2<double>* __restrict__ h, const c2<double>* __restrict__ w2,
                                              int64_t n) {
    __shared__ c2<double> cs[kBand];
    __shared__ double scale;
    if (threadIdx.x < nl) cs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ww = w2->x;
        double b2 = ww;
        for (int i = 0; i < nl; ++i) b2 -= cs[i].x * cs[i].x + cs[i].y * cs[i].y;
        const double beta = b2 > 1e-12 * ww ? sqrt(b2) : 0.0;
        scale = beta > 0.0 ? 1.0 / beta : 0.0;
        if (blockIdx.x == 0) h[nl] = c2<double>{beta, 0.0};
    }
    __syncthreads();
    const double a = scale;
    const int64_t n2 = n >> 1;
    const c2<float>* qb = Q + (int64_t)c0 * ldq;
    for (int64_t e = blockIdx.x * (int64_t)ST + threadIdx.x; e < n2; e += (int64_t)gridDim.x * ST) {
        float4 qv[kBand];
#pragma unroll
        for (int i = 0; i < kBand; ++i)
            if (i < nl) qv[i] = ld4nt(qb + (int64_t)i * ldq + 2 * e);
        const float4 wv = ld4(w + 2 * e);
        double ax = wv.x, ay = wv.y, bx = wv.z, by = wv.w;
#pragma unroll
        for (int i = 0; i < kBand; ++i) {
            if (i < nl) {
                const float4 q = qv[i];
                const c2<double> c = cs[i];
                ax -= (double)q.x * c.x - (double)q.y * c.y;
                ay -= (double)q.x * c.y + (double)q.y * c.x;
                bx -= (double)q.z * c.x - (double)q.w * c.y;
                by -= (double)q.z * c.y + (double)q.w * c.x;
            }
        }
        const float4 o = {(float)(a * ax), (float)(a * ay), (float)(a * bx), (float)(a * by)};
        *reinterpret_cast<float4*>(out + 2 * e) = o;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        double2 acc = ld2(w + n - 1);
        for (int i = 0; i < nl; ++i) {
            const double2 q = ld2(Q + (int64_t)(c0 + i) * ldq + n - 1);
            const c2<double> c = cs[i];
            acc.x -= q.x * c.x - q.y * c.y;
            acc.y -= q.x * c.y + q.y * c.x;
        }
        out[n - 1] = c2<float>{(float)(a * acc.x), (float)(a * acc.y)};
    }
}

// x += sum_{i<k} Q_i y_i (FP64 accumulation, k <= kMaxBasis)
__global__ __launch_bounds__(ST) void k_scomb(c2<double>* __restrict__ x, const c2<float>* __restrict__ Q, int64_t ldq,
                                              int k, const c2<double>* __restrict__ y, int64_t n) {
    __shared__ c2<double> cs[kMaxBasis];
    for (int i = threadIdx.x; i < k; i += ST) cs[i] = y[i];
    __syncthreads();
    // basis rows QU at a time, their (non-temporal) loads in flight together;
    // the sum runs over i in order
    constexpr int QU = 4;
    for (int64_t e = blockIdx.x * (int64_t)ST + threadIdx.x; e < n; e += (int64_t)gridDim.x * ST) {
        double2 acc = ld2(x + e);
        auto add = [&](const c2<float> a, int i) {
            const c2<double> c = cs[i];
            acc.x += (double)a.x * c.x - (double)a.y * c.y;
            acc.y += (double)a.x * c.y + (double)a.y * c.x;
        };
        int i = 0;
        for (; i + QU <= k; i += QU) {
            c2<float> a[QU];
#pragma unroll
            for (int u = 0; u < QU; ++u) a[u] = ldf_nt(Q + (int64_t)(i + u) * ldq + e);
#pragma unroll
            for (int u = 0; u < QU; ++u) add(a[u], i + u);
        }
        for (; i < k; ++i) add(ldf_nt(Q + (int64_t)i * ldq + e), i);
        x[e] = c2<double>{acc.x, acc.y};
    }
}

template <typename TV>
__global__ __launch_bounds__(ST) void k_snorm2(const c2<TV>* __restrict__ v, int64_t n, double* __restrict__ npart,
                                               int G) {
    __shared__ double sh[ST / 64];
    double acc[1] = {0.0};
    for (int64_t e = blockIdx.x * (int64_t)ST + threadIdx.x; e < n; e += (int64_t)G * ST) {
        const double2 a = ld2(v + e);
        acc[0] += a.x * a.x + a.y * a.y;
    }
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) npart[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(ST) void k_ssum(const double* __restrict__ p, int G, double* __restrict__ out) {
    __shared__ double sh[ST / 64];
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < G; b += ST) acc[0] += p[b];
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) out[0] = acc[0];
}

// q = a * r (complex64 from FP64), or from complex64 in place
template <typename TI>
__global__ void k_sscale(c2<float>* __restrict__ q, double a, const c2<TI>* __restrict__ r, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = ld2(r + e);
        q[e] = c2<float>{(float)(a * v.x), (float)(a * v.y)};
    }
}

// r = b - t
__global__ void k_ssub(c2<double>* __restrict__ r, const c2<double>* __restrict__ b, const c2<double>* __restrict__ t,
                       int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        r[e] = c2<double>{b[e].x - t[e].x, b[e].y - t[e].y};
}

struct SWs {
    c2<float>* Q = nullptr;      // (m + SM + 1) columns of ldq complex64
    c2<double>* part = nullptr;  // k_sdot partials
    c2<double>* C = nullptr;     // (m + SM + 1) x SM block coefficients (row-major), then SM x SM Gram
    c2<double>* G2 = nullptr;    // SM x SM
    c2<double>* R = nullptr;     // 4 SM x SM: R1, R1inv, R2, R2inv
    c2<double>* y = nullptr;     // kMaxBasis solution coefficients
    int* nok = nullptr;          // 3: accepted columns of the two factors, second-pass skip flag
    double* npart = nullptr;
    double* nout = nullptr;
    c2<float>* wt = nullptr;     // n: a product's output before its local orthogonalisation
    c2<double>* lpart = nullptr; // (kBand + 1) x SG local dot partials
    c2<double>* hloc = nullptr;  // SM x (SL + 1) local Hessenberg columns of a block
    c2<double>* w2 = nullptr;    // |w|^2 of the product being orthogonalised
    void* hp = nullptr;          // pinned host mirror
    size_t hp_bytes = 0;
    hipEvent_t ev = nullptr;     // behind each block's copies to the mirror
    void free_all() {
        void* p[] = {Q, part, C, G2, R, y, nok, npart, nout, wt, lpart, hloc, w2};
        for (void* q : p)
            if (q) (void)hipFree(q);
        if (hp) (void)hipHostFree(hp);
        if (ev) (void)hipEventDestroy(ev);
    }
};

struct SCtx {
    Op* h;
    hipStream_t s;
    int64_t n;
    int64_t ldq;
    int transp;
    int G;               // reduction blocks
    int m;               // most columns multiplied per cycle
    SWs w;
};

int s_grid(int64_t n) {
    int64_t g = (n + ST * 4 - 1) / (ST * 4);
    return (int)(g < 1 ? 1 : (g > SG ? SG : g));
}

template <typename TV> int s_norm(SCtx& c, const c2<TV>* v, double* out) {
    hipLaunchKernelGGL(k_snorm2<TV>, dim3(c.G), dim3(ST), 0, c.s, v, c.n, c.w.npart, c.G);
    hipLaunchKernelGGL(k_ssum, dim3(1), dim3(ST), 0, c.s, c.w.npart, c.G, c.w.nout);
    MARIE_HIP(hipGetLastError());
    double v2 = 0.0;
    MARIE_HIP(hipMemcpyAsync(&v2, c.w.nout, sizeof(double), hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipStreamSynchronize(c.s));
    *out = std::sqrt(v2);
    return OK;
}

// C (P x SM, row-major, columns past Qn zero) = Q[v0 : v0 + P)^H Q[w0 : w0 + Qn)
int s_dot(SCtx& c, int v0, int P, int w0, int Qn, c2<double>* C, const int* skip = nullptr) {
    const int ptiles = (P + STP - 1) / STP;
    // about 1024 workgroups over the column tiles
    const int G = std::max(64, std::min(SGD, SGD / ptiles));
    const dim3 grid(ptiles, G);
    hipLaunchKernelGGL(k_sdot, grid, dim3(ST), 0, c.s, c.w.Q + (int64_t)v0 * c.ldq, P, c.w.Q + (int64_t)w0 * c.ldq,
                       Qn, c.ldq, c.n, c.w.part, G, skip);
    hipLaunchKernelGGL(k_sdot_reduce, dim3((P * SM + ST / 64 - 1) / (ST / 64)), dim3(ST), 0, c.s, c.w.part, P, Qn,
                       G, C, skip);
    MARIE_HIP(hipGetLastError());
    return OK;
}

// H columns of one accepted block: k basis columns before the block, nk
// accepted products. The block's local Arnoldi gives A b_j = sum_{l<=j+1}
// Hl[j][l] b_l (b_0 = q_{k-1}, column-major ldl per j); the block pass gives
// b_{j+1} = Q C[:, j] + Q_new Rc[:, j] (C: k x SM host, row-major; Rc = R2 R1
// column-major SM x SM), so in [Q Q_new] coordinates
//   A q_{k-1} = sum_l Hl[0][l] T_l,   T_0 = e_{k-1},  T_l = [C[:, l-1]; Rc[:, l-1]]
//   A q_new_j = (A b_{j+1} - sum_{i<k} C[i][j] A q_i - sum_{i<j} Rc[i][j] A q_new_i) / Rc[j][j].
void s_hessenberg(dense::Mat& H, int ldh, int k, int nk, const std::vector<cd>& Hl, int ldl,
                  const std::vector<cd>& C, const std::vector<cd>& Rc) {
    const int kk = k + nk;
    auto tcoord = [&](int row, int l) -> cd {
        if (l == 0) return row == k - 1 ? cd(1, 0) : cd(0, 0);
        return row < k ? C[(size_t)row * SM + (l - 1)] : Rc[(size_t)(l - 1) * SM + (row - k)];
    };
    // A b_j in [Q Q_new] coordinates (j <= nk - 1 uses T_0 .. T_{j+1}, T_nk at most)
    auto ab = [&](int j, std::vector<cd>& a) {
        a.assign(kk, cd(0, 0));
        for (int l = 0; l <= j + 1; ++l) {
            const cd hl = Hl[(size_t)j * ldl + l];
            if (hl == cd(0, 0)) continue;
            for (int r = 0; r < kk; ++r) a[r] += hl * tcoord(r, l);
        }
    };
    std::vector<cd> a;
    ab(0, a);
    for (int r = 0; r < ldh; ++r) H[(size_t)(k - 1) * ldh + r] = r < kk ? a[r] : cd(0, 0);
    for (int j = 0; j + 1 < nk; ++j) {
        ab(j + 1, a);
        for (int i = 0; i < k; ++i) {
            const cd cij = C[(size_t)i * SM + j];
            if (cij == cd(0, 0)) continue;
            for (int r = 0; r < kk; ++r) a[r] -= cij * H[(size_t)i * ldh + r];
        }
        for (int i = 0; i < j; ++i) {
            const cd rij = Rc[(size_t)j * SM + i];
            for (int r = 0; r < kk; ++r) a[r] -= rij * H[(size_t)(k + i) * ldh + r];
        }
        const cd rjj = Rc[(size_t)j * SM + j];
        for (int r = 0; r < ldh; ++r) H[(size_t)(k + j) * ldh + r] = r < kk ? a[r] / rjj : cd(0, 0);
    }
}

// The cycle's least squares min || beta e_1 - H y || by Givens rotations,
// one Hessenberg column at a time (H is upper Hessenberg: the coordinates of
// p_{j+1} reach row k + j only, R being upper triangular).
struct SGivens {
    int ld;
    std::vector<double> cs;
    std::vector<cd> sn, g, R;    // R: rotated columns, column-major ld x m
    SGivens(int ld_, int m, double beta) : ld(ld_), cs(m, 1.0), sn(m, cd(0, 0)), g(m + 1, cd(0, 0)),
                                           R((size_t)ld_ * m, cd(0, 0)) {
        g[0] = cd(beta, 0);
    }
    // rotates column `col` (rows 0..col+1 of H) in; returns |residual| after it
    double add(const dense::Mat& H, int ldh, int col) {
        std::vector<cd> h(col + 2);
        for (int i = 0; i <= col + 1; ++i) h[i] = H[(size_t)col * ldh + i];
        for (int i = 0; i < col; ++i) {
            const cd a = h[i], b = h[i + 1];
            h[i] = cs[i] * a + sn[i] * b;
            h[i + 1] = -std::conj(sn[i]) * a + cs[i] * b;
        }
        const cd a = h[col], b = h[col + 1];
        const double na = std::abs(a), nb = std::abs(b);
        if (nb == 0.0) {
            cs[col] = 1.0;
            sn[col] = cd(0, 0);
        } else if (na == 0.0) {
            cs[col] = 0.0;
            sn[col] = std::conj(b) / nb;
        } else {
            const double tt = std::hypot(na, nb);
            cs[col] = na / tt;
            sn[col] = (a / na) * std::conj(b) / tt;
        }
        h[col] = cs[col] * a + sn[col] * b;
        const cd gk = g[col];
        g[col] = cs[col] * gk;
        g[col + 1] = -std::conj(sn[col]) * gk;
        for (int i = 0; i <= col; ++i) R[(size_t)col * ld + i] = h[i];
        return std::abs(g[col + 1]);
    }
    // y = R^{-1} g over the first ncol columns; false on a zero pivot
    bool solve(int ncol, std::vector<cd>& y) const {
        y.assign(ncol, cd(0, 0));
        for (int i = ncol - 1; i >= 0; --i) {
            cd v = g[i];
            for (int j = i + 1; j < ncol; ++j) v -= R[(size_t)j * ld + i] * y[j];
            const cd d = R[(size_t)i * ld + i];
            if (d == cd(0, 0)) return false;
            y[i] = v / d;
        }
        return true;
    }
};

// The block's sb products and their local Arnoldi: b_0 = q_{k-1} = Q[k-1],
// w = A b_j, h = B^H w, Q[k + j] = b_{j+1} = (w - B h) / ||.|| with B = [b_{l0}
// .. b_j], the last min(j + 1, kBand) columns (local Hessenberg column j of
// the block in hloc, on the device; its other entries zero).
int s_products(SCtx& c, int k, int sb) {
    const size_t ldq = (size_t)c.ldq;
    if (kBand < SL)
        MARIE_HIP(hipMemsetAsync(c.w.hloc, 0, (size_t)SM * (SL + 1) * sizeof(c2<double>), c.s));
    for (int j = 0; j < sb; ++j) {
        const c2<float>* in = c.w.Q + (size_t)(k - 1 + j) * ldq;
        c2<float>* out = c.w.Q + (size_t)(k + j) * ldq;
        const int st = apply(c.h, OP_AN, c.transp, PREC_F32, in, c.w.wt, 1, c.s);
        if (st) return st;
        // the product is orthogonalised against the block's last nl columns
        // b_{l0} .. b_j (all of them while j + 1 <= kBand)
        const int nl = std::min(j + 1, kBand), l0 = j + 1 - nl;
        c2<double>* hj = c.w.hloc + (size_t)j * (SL + 1) + l0;
        // a quarter of the combination's blocks: each block's 19-row LDS sum
        // then covers four times the rows
        const int gd = std::max(1, c.G / 4);
        hipLaunchKernelGGL(k_ldot, dim3(gd), dim3(ST), 0, c.s, c.w.Q, c.ldq, k - 1 + l0, nl, c.w.wt, c.n,
                           c.w.lpart, gd);
        hipLaunchKernelGGL(k_lsum, dim3(nl + 1), dim3(ST), 0, c.s, c.w.lpart, gd, nl, hj, c.w.w2);
        hipLaunchKernelGGL(k_lcomb, dim3(c.G), dim3(ST), 0, c.s, out, c.w.wt, c.w.Q, c.ldq, k - 1 + l0, nl, hj,
                           c.w.w2, c.n);
        MARIE_HIP(hipGetLastError());
    }
    return OK;
}

// Block Gram-Schmidt and CholeskyQR2 of the block at column k, then its
// small matrices (C rows 0..k-1, R1, R1inv, R2, R2inv, accepted counts) to
// the pinned mirror and an event behind them.
int s_passes(SCtx& c, int k, int sb, hipEvent_t ev) {
    c2<double>* R1 = c.w.R;
    c2<double>* R1i = c.w.R + SM * SM;
    c2<double>* R2 = c.w.R + 2 * SM * SM;
    c2<double>* R2i = c.w.R + 3 * SM * SM;
    // [C; G0] = [Q P]^H P, R1 = chol(G0 - C^H C), P = (P - Q C) R1^{-1}
    int st = s_dot(c, 0, k + sb, k, sb, c.w.C);
    if (st) return st;
    int* well = c.w.nok + 2;
    hipLaunchKernelGGL(k_schol, dim3(1), dim3(ST), 0, c.s, c.w.C, k, sb, 1, 1e-10, R1, R1i, c.w.nok,
                       (const int*)nullptr, well);
    const unsigned gu = (unsigned)((c.n + ST - 1) / ST);
    hipLaunchKernelGGL(k_supdate, dim3(gu), dim3(ST), 0, c.s, c.w.Q, c.ldq, k, sb, c.w.C, R1i, 1, c.n,
                       (const int*)nullptr);
    MARIE_HIP(hipGetLastError());
    // R2 = chol(P^H P), P = P R2^{-1} (the Gram's rows start at row 0 of G2);
    // its kernels return at once (R2 = I) when the first factor kept the
    // block well apart from the basis
    st = s_dot(c, k, sb, k, sb, c.w.G2, well);
    if (st) return st;
    hipLaunchKernelGGL(k_schol, dim3(1), dim3(ST), 0, c.s, c.w.G2, 0, sb, 0, 1e-10, R2, R2i, c.w.nok + 1,
                       (const int*)well, (int*)nullptr);
    hipLaunchKernelGGL(k_supdate, dim3(gu), dim3(ST), 0, c.s, c.w.Q, c.ldq, k, sb, c.w.C, R2i, 0, c.n,
                       (const int*)well);
    MARIE_HIP(hipGetLastError());
    unsigned char* hp = reinterpret_cast<unsigned char*>(c.w.hp);
    const size_t cbytes = (size_t)k * SM * sizeof(c2<double>);
    const size_t rbytes = (size_t)4 * SM * SM * sizeof(c2<double>);
    const size_t lbytes = (size_t)SM * (SL + 1) * sizeof(c2<double>);
    MARIE_HIP(hipMemcpyAsync(hp, c.w.C, cbytes, hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipMemcpyAsync(hp + cbytes, c.w.R, rbytes, hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipMemcpyAsync(hp + cbytes + rbytes, c.w.hloc, lbytes, hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipMemcpyAsync(hp + cbytes + rbytes + lbytes, c.w.nok, 2 * sizeof(int), hipMemcpyDeviceToHost, c.s));
    MARIE_HIP(hipEventRecord(ev, c.s));
    return OK;
}

// products the recent convergence rate predicts are still needed after a
// residual `res` (0 when the rate is not known yet)
int s_need(const std::vector<double>& hist, double res, double abstol) {
    if ((int)hist.size() <= SM) return 0;
    const double rho = std::pow(hist.back() / hist[hist.size() - 1 - SM], 1.0 / SM);
    if (!(rho > 0.0 && rho < 1.0)) return 0;
    if (res < abstol) return -1;
    return (int)std::ceil(std::log(abstol / res) / std::log(rho));
}

// One cycle on the FP64 residual r (norm beta): the correction is added to x.
// Stops when the least-squares residual is below abstol or m products are
// done. Appends the least-squares residual after every accepted product.
//
// While the host turns block j's small matrices into Hessenberg columns and
// Givens rotations, the device already runs block j+1's products (queued
// behind block j's passes unless the convergence rate predicts that block j
// ends the cycle); they are redone when block j stops early in its CholeskyQR.
int s_cycle(SCtx& c, c2<double>* x, const c2<double>* r, double beta, double abstol, int* kout,
            std::vector<double>& resv) {
    const int m = c.m;
    const int ldh = m + SM + 1;
    hipLaunchKernelGGL(k_sscale<double>, dim3(c.G), dim3(ST), 0, c.s, c.w.Q, 1.0 / beta, r, c.n);
    MARIE_HIP(hipGetLastError());
    dense::Mat H((size_t)ldh * (m + SM), cd(0, 0));
    SGivens gv(ldh, m + SM, beta);
    std::vector<cd> y, Ch, R4((size_t)4 * SM * SM), Hl((size_t)SM * (SL + 1));
    std::vector<double> hist;
    hist.push_back(beta);
    int k = 1;          // basis columns; q_{k-1} not yet multiplied
    int sb = std::min(SM, m);
    bool queued = false;   // the current block's products are already queued
    int ncol = 0;
    double res = beta;
    while (k - 1 < m) {
        int st = OK;
        if (!queued) st = s_products(c, k, sb);
        if (st) return st;
        st = s_passes(c, k, sb, c.w.ev);
        if (st) return st;
        // queue the next block unless this one is predicted to end the cycle
        int kn = k + sb, sbn = 0;
        if (kn - 1 < m) {
            const int need_now = s_need(hist, hist.back(), abstol);
            if (need_now == 0 || need_now > sb) {
                const int need_after = need_now == 0 ? SM : need_now - sb + 1;
                sbn = std::max(1, std::min(std::min(SM, need_after), m - (kn - 1)));
                st = s_products(c, kn, sbn);
                if (st) return st;
            }
        }
        MARIE_HIP(hipEventSynchronize(c.w.ev));
        unsigned char* hp = reinterpret_cast<unsigned char*>(c.w.hp);
        const size_t cbytes = (size_t)k * SM * sizeof(c2<double>);
        const size_t rbytes = (size_t)4 * SM * SM * sizeof(c2<double>);
        const size_t lbytes = (size_t)SM * (SL + 1) * sizeof(c2<double>);
        const c2<double>* hc = reinterpret_cast<const c2<double>*>(hp);
        const c2<double>* hr = reinterpret_cast<const c2<double>*>(hp + cbytes);
        const c2<double>* hl = reinterpret_cast<const c2<double>*>(hp + cbytes + rbytes);
        const int* hn = reinterpret_cast<const int*>(hp + cbytes + rbytes + lbytes);
        const int nk = std::min(hn[0], hn[1]);
        if (nk == 0) break;      // the next direction lies in the basis: the space is exhausted
        Ch.resize((size_t)k * SM);
        for (size_t i = 0; i < (size_t)k * SM; ++i) Ch[i] = cd(hc[i].x, hc[i].y);
        for (size_t i = 0; i < R4.size(); ++i) R4[i] = cd(hr[i].x, hr[i].y);
        // Rc = R2 R1 (upper triangular, leading nk x nk)
        std::vector<cd> Rc((size_t)SM * SM, cd(0, 0));
        for (int d = 0; d < nk; ++d)
            for (int cc = 0; cc <= d; ++cc) {
                cd v(0, 0);
                for (int i = cc; i <= d; ++i) v += R4[2 * SM * SM + (size_t)i * SM + cc] * R4[(size_t)d * SM + i];
                Rc[(size_t)d * SM + cc] = v;
            }
        for (size_t i = 0; i < Hl.size(); ++i) Hl[i] = cd(hl[i].x, hl[i].y);
        s_hessenberg(H, ldh, k, nk, Hl, SL + 1, Ch, Rc);
        const int k0 = k;
        // least-squares residual after each accepted product (Hessenberg
        // columns k0-1 .. k0+nk-2); a block with a non-finite column is dropped
        bool finite = true;
        for (int col = k0 - 1; col < k0 - 1 + nk && finite; ++col)
            for (int rr = 0; rr <= col + 1; ++rr)
                finite = finite && std::isfinite(H[(size_t)col * ldh + rr].real()) &&
                         std::isfinite(H[(size_t)col * ldh + rr].imag());
        if (!finite) break;
        for (int col = k0 - 1; col < k0 - 1 + nk; ++col) {
            const double rr = gv.add(H, ldh, col);
            hist.push_back(rr);
            resv.push_back(rr);
        }
        k += nk;
        res = hist.back();
        ncol = k - 1;
        if (res < abstol) break;
        if (nk == sb && sbn > 0) {
            // the queued block starts where this one ended
            sb = sbn;
            queued = true;
        } else {
            // this block stopped early (or nothing was queued): size the next
            // one by the rate and queue it from the last accepted column
            const int need = s_need(hist, res, abstol);
            const int stopped = nk < sb ? nk : SM;
            sb = need > 0 ? std::min(SM, need + 1) : SM;
            sb = std::max(1, std::min(std::min(sb, stopped), m - (k - 1)));
            queued = false;
        }
    }
    // products queued past the cycle's end only wrote basis columns past ncol
    if (ncol == 0) {
        MARIE_HIP(hipStreamSynchronize(c.s));
        *kout = 0;
        return OK;
    }
    if (!gv.solve(ncol, y)) {
        set_error("the s-step cycle's least-squares problem is singular");
        return ERR_NONFINITE;
    }
    std::vector<c2<double>> yh(ncol);
    for (int i = 0; i < ncol; ++i) yh[i] = c2<double>{y[i].real(), y[i].imag()};
    MARIE_HIP(hipMemcpyAsync(c.w.y, yh.data(), ncol * sizeof(c2<double>), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(k_scomb, dim3(c.G), dim3(ST), 0, c.s, x, c.w.Q, c.ldq, ncol, c.w.y, c.n);
    MARIE_HIP(hipGetLastError());
    MARIE_HIP(hipStreamSynchronize(c.s));
    *kout = ncol;
    return OK;
}

}  // namespace

// Mode 3 of marie_vie_gmres_device: refinement cycles of s-step GMRES with
// complex64 products. r holds b - A x on entry (FP64, norm rnorm), t is FP64
// scratch; the inner-target rule is that of the other refined modes.
int sstep_refine(Op* h, hipStream_t s, int transp, const c2<double>* b, c2<double>* x, c2<double>* r, c2<double>* t,
                 double bnorm, double rnorm, int m, double tol, int maxit, double inner_tol,
                 std::vector<double>& resvec, int* p_out, int* outit_out) {
    SCtx c;
    c.h = h;
    c.s = s;
    c.n = h->nidx;
    c.ldq = (c.n + 1) & ~(int64_t)1;
    c.transp = transp;
    c.G = s_grid(c.n);
    c.m = m;
    if (m + SM + 1 > kMaxBasis) {
        set_error("the s-step cycle takes at most " + std::to_string(kMaxBasis - SM - 1) + " products");
        return ERR_ARG;
    }
    auto A = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the s-step workspace");
            return ERR_ALLOC;
        }
        return OK;
    };
    const int ncolmax = m + SM + 1;
    int st = A((void**)&c.w.Q, (size_t)ncolmax * c.ldq * sizeof(c2<float>));
    if (!st) st = A((void**)&c.w.part, (size_t)ncolmax * SM * SGD * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.C, (size_t)ncolmax * SM * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.G2, (size_t)SM * SM * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.R, (size_t)4 * SM * SM * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.y, (size_t)kMaxBasis * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.nok, 3 * sizeof(int));
    if (!st) st = A((void**)&c.w.npart, SG * sizeof(double));
    if (!st) st = A((void**)&c.w.nout, sizeof(double));
    if (!st) st = A((void**)&c.w.wt, (size_t)c.ldq * sizeof(c2<float>));
    if (!st) st = A((void**)&c.w.lpart, (size_t)(kBand + 1) * SG * sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.w2, sizeof(c2<double>));
    if (!st) st = A((void**)&c.w.hloc, (size_t)SM * (SL + 1) * sizeof(c2<double>));
    c.w.hp_bytes = ((size_t)ncolmax * SM + 4 * SM * SM + SM * (SL + 1)) * sizeof(c2<double>) + 2 * sizeof(int);
    if (!st && hipHostMalloc(&c.w.hp, c.w.hp_bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c.w.hp = nullptr;
        set_error("pinned host allocation failed");
        st = ERR_ALLOC;
    }
    if (!st && hipEventCreateWithFlags(&c.w.ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        c.w.ev = nullptr;
        set_error("hipEventCreate failed");
        st = ERR_HIP;
    }
    if (st) {
        c.w.free_all();
        return st;
    }
    int outit = 0, p = 0;
    double rcur = rnorm;
    while (outit < maxit && rcur > tol * bnorm) {
        ++outit;
        double rel = 0.9 * tol * bnorm / rcur;
        if (inner_tol > rel) rel = inner_tol;
        if (rel > 0.5) rel = 0.5;
        std::vector<double> rv;
        st = s_cycle(c, x, r, rcur, rel * rcur, &p, rv);
        if (st) break;
        // r = b - A64 x
        st = apply(h, OP_AN, transp, PREC_F64, x, t, 1, s);
        if (st) break;
        hipLaunchKernelGGL(k_ssub, dim3(c.G), dim3(ST), 0, s, r, b, t, c.n);
        if (hipGetLastError() != hipSuccess) {
            set_error("kernel launch failed");
            st = ERR_HIP;
            break;
        }
        st = s_norm<double>(c, r, &rcur);
        if (st) break;
        for (double v : rv) resvec.push_back(v / bnorm);
        resvec.push_back(rcur / bnorm);
        if (p == 0) break;      // no progress possible in complex64
    }
    (void)hipStreamSynchronize(s);
    c.w.free_all();
    *p_out = p;
    *outit_out = outit;
    return st;
}

}  // namespace marie
