// FFT-circulant matvec of the JVIE N operator on gfx950.
//
// y = Gram*Mr*x - Mc*crop(IFFT(fN (x) FFT(pad(x))))  (and the adjoint / N / L
// variants). Five kernels, each a batched line FFT staged in LDS:
//   k_fwd_x : gather compact x through the position map, zero-pad to Px,
//             forward FFT along x              A[c][k][j][0:Px]   (j<M, k<N)
//   k_fwd_y : zero-pad along y to Py, forward  B[c][k][0:Py][i]   (k<N)
//   k_z     : per (i,j) column: forward z FFT (zero-padded to Pz) of the three
//             components, the symmetric 6-component tensor product with fN,
//             inverse z FFT, keep k<N           B (in place)
//   k_inv_y : inverse along y, keep j<M        A
//   k_inv_x : inverse along x, keep i<L, fused material epilogue and scatter
//             to the compact output through the position map.
// Only the non-zero part of each padded grid is ever stored: the forward
// passes skip the zero blocks and the inverse passes drop what the crop
// discards, so the intermediates are 2n (A) and 4n (B) complex per component
// instead of 8n.
#include "vie_handle.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace marie {

constexpr int NT = 256;

// -------------------------------------------------------------------- kernels
template <typename T, int EMAX, int RSET>
__global__ __launch_bounds__(NT) void k_fwd_x(const c2<T>* __restrict__ x, c2<T>* __restrict__ A,
                                              const int* __restrict__ pos, const c2<T>* __restrict__ Mc,
                                              const c2<T>* __restrict__ tw, AxisPlan pl, int L, int M,
                                              int N, int lg, int64_t nidx, int conj_in) {
    extern __shared__ unsigned char smem_[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_);
    const int TX = 1 << lg;
    const int P = pl.P;
    const int comp = blockIdx.y;
    const int c = comp % 3, rhs = comp / 3;
    const int64_t nline = (int64_t)M * N;
    const int64_t nv = nline * L;
    const int64_t line0 = (int64_t)blockIdx.x * TX;
    const c2<T>* xr = x + rhs * nidx;
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e / P, i = e - b * P;
        c2<T> v = {T(0), T(0)};
        const int64_t line = line0 + b;
        if (i < L && line < nline) {
            const int64_t vox = i + (int64_t)L * line;
            const int p = pos[vox + c * nv];
            if (p >= 0) {
                v = xr[p];
                if (conj_in) v = cmul(Mc[vox], cconj(v));
            }
        }
        buf[i * TX + b] = v;
    }
    __syncthreads();
    tile_fft<T, EMAX, false, RSET>(buf, tw, pl, lg);
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e / P, i = e - b * P;
        const int64_t line = line0 + b;
        if (line < nline) A[((int64_t)comp * nline + line) * P + i] = buf[i * TX + b];
    }
}

// Strided-axis pass: lines along an axis of stride `ls` (elements), tile of TX
// consecutive x positions. `nin` inputs are read (rest zero), `nout` outputs kept.
template <typename T, int EMAX, bool INV, int RSET>
__global__ __launch_bounds__(NT) void k_strided(const c2<T>* __restrict__ src, c2<T>* __restrict__ dst,
                                                const c2<T>* __restrict__ tw, AxisPlan pl, int Px,
                                                int64_t s_ls, int64_t s_plane, int64_t d_ls,
                                                int64_t d_plane, int nin, int nout, int lg,
                                                int64_t s_comp, int64_t d_comp) {
    extern __shared__ unsigned char smem_[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_);
    const int TX = 1 << lg;
    const int msk = TX - 1;
    const int P = pl.P;
    const int i0 = blockIdx.x * TX;
    const int64_t plane = blockIdx.y;
    const int comp = blockIdx.z;
    const c2<T>* s = src + comp * s_comp + plane * s_plane + i0;
    c2<T>* d = dst + comp * d_comp + plane * d_plane + i0;
    const int nx = min(TX, Px - i0);
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e & msk, p = e >> lg;
        c2<T> v = {T(0), T(0)};
        if (p < nin && b < nx) v = s[p * s_ls + b];
        buf[e] = v;
    }
    __syncthreads();
    tile_fft<T, EMAX, INV, RSET>(buf, tw, pl, lg);
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e & msk, p = e >> lg;
        if (p < nout && b < nx) d[p * d_ls + b] = buf[e];
    }
}

// Fused z pass: forward z FFT of the three components, tensor product, inverse.
template <typename T, int EMAX, int RSET>
__global__ __launch_bounds__(NT) void k_z(c2<T>* __restrict__ B, const c2<T>* __restrict__ fN,
                                          const c2<T>* __restrict__ tw, AxisPlan pl, int Px, int Py,
                                          int N, int lg) {
    extern __shared__ unsigned char smem_[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_);
    const int TX = 1 << lg;
    const int msk = TX - 1;
    const int P = pl.P;
    const int i0 = blockIdx.x * TX;
    const int j = blockIdx.y;
    const int rhs = blockIdx.z;
    const int64_t plane = (int64_t)Py * Px;
    const int64_t cstride = plane * N;
    const int64_t fstride = plane * P;
    const int nx = min(TX, Px - i0);
    const int tot = TX * P;
    c2<T>* base = B + (int64_t)rhs * 3 * cstride + (int64_t)j * Px + i0;
    c2<T> X[3][EMAX];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int e = threadIdx.x; e < tot; e += NT) {
            const int b = e & msk, k = e >> lg;
            c2<T> v = {T(0), T(0)};
            if (k < N && b < nx) v = base[c * cstride + k * plane + b];
            buf[e] = v;
        }
        __syncthreads();
        tile_fft<T, EMAX, false, RSET>(buf, tw, pl, lg);
#pragma unroll
        for (int q = 0; q < EMAX; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e < tot) X[c][q] = buf[e];
        }
        __syncthreads();
    }
    const c2<T>* f = fN + (int64_t)j * Px + i0;
#pragma unroll
    for (int q = 0; q < EMAX; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < tot) {
            const int b = e & msk, k = e >> lg;
            if (b < nx) {
                const int64_t o = k * plane + b;
                const c2<T> f0 = f[o], f1 = f[o + fstride], f2 = f[o + 2 * fstride];
                const c2<T> f3 = f[o + 3 * fstride], f4 = f[o + 4 * fstride], f5 = f[o + 5 * fstride];
                const c2<T> x0 = X[0][q], x1 = X[1][q], x2 = X[2][q];
                X[0][q] = cadd(cadd(cmul(f0, x0), cmul(f1, x1)), cmul(f2, x2));
                X[1][q] = cadd(cadd(cmul(f1, x0), cmul(f3, x1)), cmul(f4, x2));
                X[2][q] = cadd(cadd(cmul(f2, x0), cmul(f4, x1)), cmul(f5, x2));
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int q = 0; q < EMAX; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e < tot) buf[e] = X[a][q];
        }
        __syncthreads();
        tile_fft<T, EMAX, true, RSET>(buf, tw, pl, lg);
        for (int e = threadIdx.x; e < tot; e += NT) {
            const int b = e & msk, k = e >> lg;
            if (k < N && b < nx) base[a * cstride + k * plane + b] = buf[e];
        }
        __syncthreads();
    }
}

// Inverse x pass with the crop, the 1/(Px Py Pz) scale and the material epilogue.
template <typename T, int EMAX, int RSET>
__global__ __launch_bounds__(NT) void k_inv_x(const c2<T>* __restrict__ A, c2<T>* __restrict__ y,
                                              const c2<T>* __restrict__ x, const int* __restrict__ pos,
                                              const c2<T>* __restrict__ Mr, const c2<T>* __restrict__ Mc,
                                              const c2<T>* __restrict__ tw, AxisPlan pl, int L, int M,
                                              int N, int lg, int64_t nidx, int op, int transp, T Gram,
                                              T scale) {
    extern __shared__ unsigned char smem_[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_);
    const int TX = 1 << lg;
    const int P = pl.P;
    const int comp = blockIdx.y;
    const int c = comp % 3, rhs = comp / 3;
    const int64_t nline = (int64_t)M * N;
    const int64_t nv = nline * L;
    const int64_t line0 = (int64_t)blockIdx.x * TX;
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e / P, i = e - b * P;
        const int64_t line = line0 + b;
        c2<T> v = {T(0), T(0)};
        if (line < nline) v = A[((int64_t)comp * nline + line) * P + i];
        buf[i * TX + b] = v;
    }
    __syncthreads();
    tile_fft<T, EMAX, true, RSET>(buf, tw, pl, lg);
    const c2<T>* xr = x + rhs * nidx;
    c2<T>* yr = y + rhs * nidx;
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e / P, i = e - b * P;
        const int64_t line = line0 + b;
        if (i < L && line < nline) {
            const int64_t vox = i + (int64_t)L * line;
            const int p = pos[vox + c * nv];
            if (p >= 0) {
                const c2<T> o = cscale(buf[i * TX + b], scale);
                c2<T> r;
                if (op == OP_AN) {
                    const c2<T> xv = xr[p];
                    if (transp) {
                        r = csub(cscale(cmulc(xv, Mr[vox]), Gram), cconj(o));
                    } else {
                        r = csub(cscale(cmul(Mr[vox], xv), Gram), cmul(Mc[vox], o));
                    }
                } else if (op == OP_L) {
                    r = csub(o, cscale(xr[p], Gram));
                } else {
                    r = o;
                }
                yr[p] = r;
            }
        }
    }
}

// -------------------------------------------------------------------- planning
int build_plan(int P, AxisPlan* pl) {
    pl->P = P;
    pl->nst = 0;
    int r = P;
    auto push = [&](int R) {
        if (pl->nst >= kMaxStages) return false;
        pl->radix[pl->nst++] = R;
        return true;
    };
    while (r % 8 == 0) { if (!push(8)) return ERR_SIZE; r /= 8; }
    if (r % 4 == 0) { if (!push(4)) return ERR_SIZE; r /= 4; }
    if (r % 2 == 0) { if (!push(2)) return ERR_SIZE; r /= 2; }
    const int odd[] = {3, 5, 7, 11, 13};
    for (int R : odd)
        while (r % R == 0) { if (!push(R)) return ERR_SIZE; r /= R; }
    if (r != 1) {
        set_error("FFT length " + std::to_string(P) +
                  " has a prime factor above 13; choose a circulant size with factors 2,3,5,7,11,13");
        return ERR_SIZE;
    }
    return OK;
}

int upload_twiddles(Op* h) {
    const int P[3] = {h->Px, h->Py, h->Pz};
    const int tot = P[0] + P[1] + P[2];
    std::vector<double> t64(2 * (size_t)tot);
    std::vector<float> t32(2 * (size_t)tot);
    size_t o = 0;
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < P[a]; ++k, ++o) {
            // exp(-2 pi i k / P) evaluated in extended precision, rounded once
            const long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / P[a];
            const long double cr = cosl(ang), ci = -sinl(ang);
            t64[2 * o] = (double)cr;
            t64[2 * o + 1] = (double)ci;
            t32[2 * o] = (float)cr;
            t32[2 * o + 1] = (float)ci;
        }
    }
    if (!h->tw64) MARIE_HIP(hipMalloc(&h->tw64, t64.size() * sizeof(double)));
    if (!h->tw32) MARIE_HIP(hipMalloc(&h->tw32, t32.size() * sizeof(float)));
    MARIE_HIP(hipMemcpy(h->tw64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
    MARIE_HIP(hipMemcpy(h->tw32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
    return OK;
}

int ensure_work(Op* h, int nrhs) {
    if (nrhs <= h->nrhs_cap) return OK;
    const size_t es = (h->precisions & PREC_F64) ? sizeof(c2<double>) : sizeof(c2<float>);
    const size_t na = (size_t)3 * nrhs * h->Px * h->M * h->N;
    const size_t nb = (size_t)3 * nrhs * h->Px * h->Py * h->N;
    if (h->A) { MARIE_HIP(hipFree(h->A)); h->A = nullptr; }
    if (h->B) { MARIE_HIP(hipFree(h->B)); h->B = nullptr; }
    h->nrhs_cap = 0;
    if (hipMalloc(&h->A, na * es) != hipSuccess || hipMalloc(&h->B, nb * es) != hipSuccess) {
        (void)hipGetLastError();
        if (h->A) (void)hipFree(h->A);
        h->A = h->B = nullptr;
        set_error("out of device memory for matvec work buffers");
        return ERR_ALLOC;
    }
    h->nrhs_cap = nrhs;
    return OK;
}

// largest power of two TX with TX*P <= cap and TX <= txmax
static int pick_lg(int P, int cap, int txmax) {
    int lg = 0;
    while ((2 << lg) * P <= cap && (2 << lg) <= txmax) ++lg;
    return lg;
}

template <typename T, int EMAX_XY, int EMAX_Z>
static int apply_t(Op* h, int op, int transp, const c2<T>* x, c2<T>* y, int nrhs, hipStream_t s) {
    const bool f64 = sizeof(T) == 8;
    const c2<T>* fN = reinterpret_cast<const c2<T>*>(f64 ? h->fN64 : h->fN32);
    const c2<T>* Mr = reinterpret_cast<const c2<T>*>(f64 ? h->Mr64 : h->Mr32);
    const c2<T>* Mc = reinterpret_cast<const c2<T>*>(f64 ? h->Mc64 : h->Mc32);
    const c2<T>* tw = reinterpret_cast<const c2<T>*>(f64 ? h->tw64 : h->tw32);
    const c2<T>* twx = tw;
    const c2<T>* twy = tw + h->Px;
    const c2<T>* twz = tw + h->Px + h->Py;
    c2<T>* A = reinterpret_cast<c2<T>*>(h->A);
    c2<T>* B = reinterpret_cast<c2<T>*>(h->B);
    const int L = h->L, M = h->M, N = h->N, Px = h->Px, Py = h->Py, Pz = h->Pz;
    const int ncomp = 3 * nrhs;
    const int capxy = EMAX_XY * NT, capz = EMAX_Z * NT;
    const int txmax = f64 ? 16 : 32;
    if (Px > capxy || Py > capxy || Pz > capz) {
        set_error("circulant side exceeds the single-tile FFT limit");
        return ERR_SIZE;
    }
    const int64_t nline = (int64_t)M * N;
    const bool tm = h->timing;
    if (tm) MARIE_HIP(hipEventRecord(h->ev[0], s));

    // forward x
    {
        const int lg = pick_lg(Px, capxy, 16);
        const dim3 g((unsigned)((nline + (1 << lg) - 1) >> lg), ncomp);
        const size_t sm = ((size_t)Px << lg) * sizeof(c2<T>);
        MARIE_RSET(h->plan[0], hipLaunchKernelGGL((k_fwd_x<T, EMAX_XY, RS>), g, dim3(NT), sm, s, x, A, h->pos, Mc, twx, h->plan[0], L,
                           M, N, lg, h->nidx, (op == OP_AN && transp) ? 1 : 0));
        MARIE_HIP(hipGetLastError());
    }
    if (tm) MARIE_HIP(hipEventRecord(h->ev[1], s));
    // forward y: A[c][k][j<M][i] -> B[c][k][j<Py][i]
    {
        const int lg = pick_lg(Py, capxy, txmax);
        const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg), N, ncomp);
        const size_t sm = ((size_t)Py << lg) * sizeof(c2<T>);
        MARIE_RSET(h->plan[1], hipLaunchKernelGGL((k_strided<T, EMAX_XY, false, RS>), g, dim3(NT), sm, s, A, B, twy, h->plan[1], Px,
                           (int64_t)Px, (int64_t)M * Px, (int64_t)Px, (int64_t)Py * Px, M, Py, lg,
                           (int64_t)N * M * Px, (int64_t)N * Py * Px));
        MARIE_HIP(hipGetLastError());
    }
    if (tm) MARIE_HIP(hipEventRecord(h->ev[2], s));
    // z: fused forward, tensor product, inverse
    {
        const int lg = pick_lg(Pz, capz, txmax);
        const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg), Py, nrhs);
        const size_t sm = ((size_t)Pz << lg) * sizeof(c2<T>);
        MARIE_RSET(h->plan[2], hipLaunchKernelGGL((k_z<T, EMAX_Z, RS>), g, dim3(NT), sm, s, B, fN, twz, h->plan[2], Px, Py, N, lg));
        MARIE_HIP(hipGetLastError());
    }
    if (tm) MARIE_HIP(hipEventRecord(h->ev[3], s));
    // inverse y: B[c][k][j<Py][i] -> A[c][k][j<M][i]
    {
        const int lg = pick_lg(Py, capxy, txmax);
        const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg), N, ncomp);
        const size_t sm = ((size_t)Py << lg) * sizeof(c2<T>);
        MARIE_RSET(h->plan[1], hipLaunchKernelGGL((k_strided<T, EMAX_XY, true, RS>), g, dim3(NT), sm, s, B, A, twy, h->plan[1], Px,
                           (int64_t)Px, (int64_t)Py * Px, (int64_t)Px, (int64_t)M * Px, Py, M, lg,
                           (int64_t)N * Py * Px, (int64_t)N * M * Px));
        MARIE_HIP(hipGetLastError());
    }
    if (tm) MARIE_HIP(hipEventRecord(h->ev[4], s));
    // inverse x + crop + epilogue
    {
        const int lg = pick_lg(Px, capxy, 16);
        const dim3 g((unsigned)((nline + (1 << lg) - 1) >> lg), ncomp);
        const size_t sm = ((size_t)Px << lg) * sizeof(c2<T>);
        const T scale = T(1.0 / ((double)Px * (double)Py * (double)Pz));
        MARIE_RSET(h->plan[0], hipLaunchKernelGGL((k_inv_x<T, EMAX_XY, RS>), g, dim3(NT), sm, s, A, y, x, h->pos, Mr, Mc, twx,
                           h->plan[0], L, M, N, lg, h->nidx, op, transp, (T)h->Gram, scale));
        MARIE_HIP(hipGetLastError());
    }
    if (tm) {
        MARIE_HIP(hipEventRecord(h->ev[5], s));
        MARIE_HIP(hipEventSynchronize(h->ev[5]));
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            MARIE_HIP(hipEventElapsedTime(&ms, h->ev[k], h->ev[k + 1]));
            h->kernel_ms[k] += ms;
        }
        h->kernel_calls += 1;
    }
    return OK;
}

int apply(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s) {
    if (!h->have_fN || !h->have_idx || (op == OP_AN && !h->have_mat)) {
        set_error("operator, index or material not set");
        return ERR_STATE;
    }
    if (!(h->precisions & prec)) {
        set_error("requested precision was not enabled at create");
        return ERR_STATE;
    }
    int st = ensure_work(h, nrhs);
    if (st) return st;
    if (prec == PREC_F64)
        return apply_t<double, 16, 8>(h, op, transp, (const c2<double>*)x, (c2<double>*)y, nrhs, s);
    return apply_t<float, 32, 16>(h, op, transp, (const c2<float>*)x, (c2<float>*)y, nrhs, s);
}

}  // namespace marie
