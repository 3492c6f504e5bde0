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
#include "vie_reg.hpp"

#include <algorithm>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>

namespace marie {

constexpr int NT = 256;

__global__ void k_to_f32_full(const c2<double>* __restrict__ a, c2<float>* __restrict__ b, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        b[e] = c2<float>{(float)a[e].x, (float)a[e].y};
}

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
    c2<T>* twl = buf + TX * P;          // the axis's twiddles, staged behind the tile
    for (int e = threadIdx.x; e < P; e += NT) twl[e] = tw[e];
    lds_barrier();
    tile_fft<T, EMAX, false, RSET>(buf, twl, pl, lg);
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
    c2<T>* twl = buf + TX * P;
    for (int e = threadIdx.x; e < P; e += NT) twl[e] = tw[e];
    lds_barrier();
    tile_fft<T, EMAX, INV, RSET>(buf, twl, pl, lg);
    for (int e = threadIdx.x; e < TX * P; e += NT) {
        const int b = e & msk, p = e >> lg;
        if (p < nout && b < nx) d[p * d_ls + b] = buf[e];
    }
}

// Non-temporal access to the z pass's work buffer in FP64, where it is read
// once and written once: the operator octant, read by up to 8 mirror tiles,
// keeps its place in the caches. (In complex64 the buffer is half the size
// and is still partly cached from the y pass that wrote it; there plain
// accesses measured faster.)
template <typename T>
__device__ __forceinline__ c2<T> ld_nt(const c2<T>* p) {
    if constexpr (sizeof(T) == 8) {
        typedef T v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return c2<T>{v.x, v.y};
    } else {
        return *p;
    }
}
template <typename T>
__device__ __forceinline__ void st_nt(c2<T>* p, c2<T> a) {
    if constexpr (sizeof(T) == 8) {
        typedef T v2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(v2{a.x, a.y}, reinterpret_cast<v2*>(p));
    } else {
        *p = a;
    }
}

// Fused z pass: forward z FFT of the three components, tensor product, inverse.
// OCT: fN is stored as its octant [Pz/2+1][Py/2+1][Px/2+1][6] (K: [..][3]). The circulant of
// each tensor component is even or odd under reflection of each axis
// (gperiodic_coeff_nop.m), so its DFT has the same parity:
// F(P-k) = s F(k) with s = -1 iff exactly one of the component's two indices is
// the reflected axis (xy: x,y; xz: x,z; yz: y,z).
template <typename T, int EMAX, int RSET, bool OCT, int PROD>
__global__ __launch_bounds__(NT) void k_z(c2<T>* __restrict__ B, const c2<T>* __restrict__ fN,
                                          const c2<T>* __restrict__ tw, AxisPlan pl, int Px, int Py,
                                          int N, int lg, int nrhs) {
    extern __shared__ unsigned char smem_[];
    c2<T>* buf = reinterpret_cast<c2<T>*>(smem_);
    const int TX = 1 << lg;
    const int msk = TX - 1;
    const int P = pl.P;
    constexpr int NC = PROD ? 3 : 6;    // operator components
    int bx, j, rhs;
    zpass_tile((int)blockIdx.x, (int)gridDim.x, (Px + TX - 1) / TX, Py, nrhs, bx, j, rhs);
    const int i0 = bx * TX;
    const int64_t plane = (int64_t)Py * Px;
    const int64_t cstride = plane * N;
    const int nx = min(TX, Px - i0);
    const int tot = TX * P;
    c2<T>* base = B + (int64_t)rhs * 3 * cstride + (int64_t)j * Px + i0;
    c2<T> X[3][EMAX];
    c2<T>* twl = buf + tot;             // the axis's twiddles, staged behind the tile
    for (int e = threadIdx.x; e < P; e += NT) twl[e] = tw[e];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        for (int e = threadIdx.x; e < tot; e += NT) {
            const int b = e & msk, k = e >> lg;
            c2<T> v = {T(0), T(0)};
            if (k < N && b < nx) v = ld_nt(base + c * cstride + k * plane + b);
            buf[e] = v;
        }
        lds_barrier();
        tile_fft<T, EMAX, false, RSET>(buf, twl, pl, lg);
#pragma unroll
        for (int q = 0; q < EMAX; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e < tot) X[c][q] = buf[e];
        }
        lds_barrier();
    }
    // fN addressing
    int64_t fplane, fstride, frow;
    int hx = 0, hy = 0, hz = 0;
    bool fy = false;
    if (OCT) {
        hx = Px / 2; hy = Py / 2; hz = P / 2;
        fplane = (int64_t)(hy + 1) * (hx + 1);
        fstride = 1;                // components interleaved per octant point
        fy = j > hy;
        frow = (int64_t)(fy ? Py - j : j) * (hx + 1);
    } else {
        fplane = plane;
        fstride = plane * P;
        frow = (int64_t)j * Px + i0;
    }
#pragma unroll
    for (int q = 0; q < EMAX; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < tot) {
            const int b = e & msk, k = e >> lg;
            if (b < nx) {
                const int i = i0 + b;
                const bool fx = OCT && i > hx, fz = OCT && k > hz;
                const int64_t o = OCT ? ((int64_t)(fz ? P - k : k) * fplane + frow + (fx ? Px - i : i)) * NC
                                      : k * fplane + frow + b;
                freq_product<T, PROD, OCT>(fN + o, fstride, fx, fy, fz, X[0][q], X[1][q], X[2][q]);
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
        lds_barrier();
        tile_fft<T, EMAX, true, RSET>(buf, twl, pl, lg);
        for (int e = threadIdx.x; e < tot; e += NT) {
            const int b = e & msk, k = e >> lg;
            if (k < N && b < nx) st_nt(base + a * cstride + k * plane + b, buf[e]);
        }
        lds_barrier();
    }
}

// ------------------------------------------------------ operator storage (octant)
// kind 0: the N operator's 6 tensor components (xx, xy, xz, yy, yz, zz), odd
// under reflection of an axis iff exactly one of the component's indices is
// that axis (gperiodic_coeff_nop.m); kind 1: the K operator's 3 vector
// components, component c odd under reflection of axis c (gperiodic_coeff_kop.m).
__device__ __forceinline__ void parity(int kind, int m, bool& ox, bool& oy, bool& oz) {
    if (kind == 0) {
        ox = (m == 1 || m == 2);
        oy = (m == 1 || m == 4);
        oz = (m == 2 || m == 4);
    } else {
        ox = m == 0;
        oy = m == 1;
        oz = m == 2;
    }
}

// max |F(k) - s F(P-k)| over all entries (per-block partial maxima), and max |F|
__global__ void k_sym_check(const c2<double>* __restrict__ F, int Px, int Py, int Pz, int kind, int nc,
                            double* __restrict__ part) {
    const int64_t tot = (int64_t)Px * Py * Pz * nc;
    double md = 0.0, ma = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Px);
        const int j = (int)((e / Px) % Py);
        const int k = (int)((e / ((int64_t)Px * Py)) % Pz);
        const int m = (int)(e / ((int64_t)Px * Py * Pz));
        const int mi = i ? Px - i : 0, mj = j ? Py - j : 0, mk = k ? Pz - k : 0;
        const c2<double> a = F[e];
        bool ox, oy, oz;
        parity(kind, m, ox, oy, oz);
        const int64_t base = (int64_t)m * Px * Py * Pz;
        const c2<double> bx = F[base + ((int64_t)k * Py + j) * Px + mi];
        const c2<double> by = F[base + ((int64_t)k * Py + mj) * Px + i];
        const c2<double> bz = F[base + ((int64_t)mk * Py + j) * Px + i];
        const double sx = ox ? -1.0 : 1.0, sy = oy ? -1.0 : 1.0, sz = oz ? -1.0 : 1.0;
        md = fmax(md, fabs(a.x - sx * bx.x) + fabs(a.y - sx * bx.y));
        md = fmax(md, fabs(a.x - sy * by.x) + fabs(a.y - sy * by.y));
        md = fmax(md, fabs(a.x - sz * bz.x) + fabs(a.y - sz * bz.y));
        ma = fmax(ma, fabs(a.x) + fabs(a.y));
    }
    __shared__ double s1[256], s2[256];
    s1[threadIdx.x] = md;
    s2[threadIdx.x] = ma;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

template <typename T>
__global__ void k_oct_compress(const c2<double>* __restrict__ F, c2<T>* __restrict__ O, int Px, int Py, int Pz,
                               int nc) {
    const int hx = Px / 2 + 1, hy = Py / 2 + 1, hz = Pz / 2 + 1;
    const int64_t tot = (int64_t)hx * hy * hz * nc;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(e % nc);
        const int64_t r = e / nc;
        const int i = (int)(r % hx);
        const int j = (int)((r / hx) % hy);
        const int k = (int)(r / ((int64_t)hx * hy));
        const c2<double> v = F[(((int64_t)m * Pz + k) * Py + j) * Px + i];
        O[e] = c2<T>{(T)v.x, (T)v.y};
    }
}

__global__ void k_oct_expand(const c2<double>* __restrict__ O, c2<double>* __restrict__ F, int Px, int Py, int Pz,
                             int kind, int nc) {
    const int hx = Px / 2, hy = Py / 2, hz = Pz / 2;
    const int64_t tot = (int64_t)Px * Py * Pz * nc;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Px);
        const int j = (int)((e / Px) % Py);
        const int k = (int)((e / ((int64_t)Px * Py)) % Pz);
        const int m = (int)(e / ((int64_t)Px * Py * Pz));
        const bool fx = i > hx, fy = j > hy, fz = k > hz;
        const int ie = fx ? Px - i : i, je = fy ? Py - j : j, ke = fz ? Pz - k : k;
        c2<double> v = O[((((int64_t)ke * (hy + 1) + je) * (hx + 1) + ie)) * nc + m];
        bool ox, oy, oz;
        parity(kind, m, ox, oy, oz);
        if ((ox && fx) != ((oy && fy) != (oz && fz))) v = c2<double>{-v.x, -v.y};
        F[e] = v;
    }
}

namespace {
struct CircSlot {
    void** d64;
    void** d32;
    bool* oct;
    bool* have;
    int nc;
};
CircSlot slot(Op* h, int kind) {
    if (kind == 0) return {&h->fN64, &h->fN32, &h->fN_oct, &h->have_fN, 6};
    return {&h->fK64, &h->fK32, &h->fK_oct, &h->have_fK, 3};
}
}  // namespace

// Takes ownership of a device buffer holding the full complex128 operator
// (kind 0: fN, 6 components; kind 1: fK, 3 components) and stores it on the
// handle: octant form when every component has the circulant's parity to
// 1e-14 of its largest entry (even sides only), full form otherwise.
int store_circ(Op* h, int kind, void* full64) {
    const CircSlot c = slot(h, kind);
    const int Px = h->Px, Py = h->Py, Pz = h->Pz;
    const size_t nf = (size_t)c.nc * Px * Py * Pz;
    hipStream_t s = h->stream;
    bool oct = false;
    if (!(Px & 1) && !(Py & 1) && !(Pz & 1)) {
        const int G = 1024;
        double* part = nullptr;
        MARIE_HIP(hipMalloc(&part, 2 * G * sizeof(double)));
        hipLaunchKernelGGL(k_sym_check, dim3(G), dim3(256), 0, s, (const c2<double>*)full64, Px, Py, Pz, kind, c.nc,
                           part);
        std::vector<double> hp(2 * G);
        MARIE_HIP(hipMemcpyAsync(hp.data(), part, 2 * G * sizeof(double), hipMemcpyDeviceToHost, s));
        MARIE_HIP(hipStreamSynchronize(s));
        MARIE_HIP(hipFree(part));
        double md = 0.0, ma = 0.0;
        for (int b = 0; b < G; ++b) { md = std::max(md, hp[2 * b]); ma = std::max(ma, hp[2 * b + 1]); }
        oct = ma > 0.0 && md <= 1e-14 * ma;
    }
    if (*c.d64) { MARIE_HIP(hipFree(*c.d64)); *c.d64 = nullptr; }
    if (*c.d32) { MARIE_HIP(hipFree(*c.d32)); *c.d32 = nullptr; }
    *c.oct = oct;
    if (oct) {
        const size_t no = (size_t)c.nc * (Px / 2 + 1) * (Py / 2 + 1) * (Pz / 2 + 1);
        if (h->precisions & PREC_F64) {
            MARIE_HIP(hipMalloc(c.d64, no * sizeof(c2<double>)));
            hipLaunchKernelGGL(k_oct_compress<double>, dim3(4096), dim3(256), 0, s, (const c2<double>*)full64,
                               (c2<double>*)*c.d64, Px, Py, Pz, c.nc);
        }
        if (h->precisions & PREC_F32) {
            MARIE_HIP(hipMalloc(c.d32, no * sizeof(c2<float>)));
            hipLaunchKernelGGL(k_oct_compress<float>, dim3(4096), dim3(256), 0, s, (const c2<double>*)full64,
                               (c2<float>*)*c.d32, Px, Py, Pz, c.nc);
        }
        MARIE_HIP(hipGetLastError());
        MARIE_HIP(hipStreamSynchronize(s));
        MARIE_HIP(hipFree(full64));
    } else {
        if (h->precisions & PREC_F32) {
            MARIE_HIP(hipMalloc(c.d32, nf * sizeof(c2<float>)));
            hipLaunchKernelGGL(k_to_f32_full, dim3(4096), dim3(256), 0, s, (const c2<double>*)full64,
                               (c2<float>*)*c.d32, (int64_t)nf);
            MARIE_HIP(hipGetLastError());
        }
        if (h->precisions & PREC_F64) {
            *c.d64 = full64;
        } else {
            MARIE_HIP(hipStreamSynchronize(s));
            MARIE_HIP(hipFree(full64));
        }
        MARIE_HIP(hipStreamSynchronize(s));
    }
    *c.have = true;
    return OK;
}

// full complex128 operator into a caller device buffer (expands the octant form)
int load_circ(Op* h, int kind, void* out64) {
    const CircSlot c = slot(h, kind);
    const size_t nf = (size_t)c.nc * h->Px * h->Py * h->Pz;
    if (!*c.d64) { set_error("no FP64 operator of this kind on the handle"); return ERR_STATE; }
    if (*c.oct) {
        hipLaunchKernelGGL(k_oct_expand, dim3(4096), dim3(256), 0, h->stream, (const c2<double>*)*c.d64,
                           (c2<double>*)out64, h->Px, h->Py, h->Pz, kind, c.nc);
        MARIE_HIP(hipGetLastError());
    } else {
        MARIE_HIP(hipMemcpyAsync(out64, *c.d64, nf * sizeof(c2<double>), hipMemcpyDeviceToDevice, h->stream));
    }
    return OK;
}

int store_fN(Op* h, void* full64) { return store_circ(h, 0, full64); }
int load_fN(Op* h, void* out64) { return load_circ(h, 0, out64); }

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
    c2<T>* twl = buf + TX * P;
    for (int e = threadIdx.x; e < P; e += NT) twl[e] = tw[e];
    lds_barrier();
    tile_fft<T, EMAX, true, RSET>(buf, twl, pl, lg);
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

// Tile variants. Each pass runs with EMAX elements per thread (so a workgroup of
// NT threads holds TX = EMAX*NT/P lines in LDS); the compiled set below is what
// the autotuner chooses from. The radix set with 11 and 13 only gets the default.
template <typename T> struct Variants;
template <> struct Variants<double> {
    static constexpr int xy[3] = {4, 8, 16};
    static constexpr int z[2] = {4, 8};
    static constexpr int txmax = 16;
};
template <> struct Variants<float> {
    static constexpr int xy[3] = {8, 16, 32};
    static constexpr int z[2] = {8, 16};
    static constexpr int txmax = 32;
};

#define MARIE_EMAX_XY(T, emax, RSv, STMT)                                                   \
    do {                                                                                    \
        if constexpr (sizeof(T) == 8) {                                                     \
            if ((emax) == 4 && RSv < 3) { constexpr int EM = 4; STMT; }                     \
            else if ((emax) == 8 && RSv < 3) { constexpr int EM = 8; STMT; }                \
            else { constexpr int EM = 16; STMT; }                                           \
        } else {                                                                            \
            if ((emax) == 8 && RSv < 3) { constexpr int EM = 8; STMT; }                     \
            else if ((emax) == 16 && RSv < 3) { constexpr int EM = 16; STMT; }              \
            else { constexpr int EM = 32; STMT; }                                           \
        }                                                                                   \
    } while (0)
#define MARIE_EMAX_Z(T, emax, RSv, STMT)                                                    \
    do {                                                                                    \
        if constexpr (sizeof(T) == 8) {                                                     \
            if ((emax) == 4 && RSv < 3) { constexpr int EM = 4; STMT; }                     \
            else { constexpr int EM = 8; STMT; }                                            \
        } else {                                                                            \
            if ((emax) == 8 && RSv < 3) { constexpr int EM = 8; STMT; }                     \
            else { constexpr int EM = 16; STMT; }                                           \
        }                                                                                   \
    } while (0)

static int effective_emax(int emax, const AxisPlan& pl, int dflt) {
    return plan_rset(pl) == RSET_ALL ? dflt : emax;
}

// The elements per thread that MARIE_EMAX_XY / MARIE_EMAX_Z dispatch for a
// requested value: the tile size is derived from this, never from the raw
// request, so that the tile always fits the instantiated kernel.
static int dispatched_emax(bool f64, bool zpass, int e, int rs) {
    if (zpass) {
        if (f64) return (e == 4 && rs < 3) ? 4 : 8;
        return (e == 8 && rs < 3) ? 8 : 16;
    }
    if (f64) return (e == 4 && rs < 3) ? 4 : (e == 8 && rs < 3) ? 8 : 16;
    return (e == 8 && rs < 3) ? 8 : (e == 16 && rs < 3) ? 16 : 32;
}

// Preconditioner material of JVIE_Solver.m:278, :297 (mv_AN called with
// 1./Mr and 1./Mr - 1), derived from Mr in FP64, rounded once for complex64.
__global__ void k_precond_material(const c2<double>* __restrict__ Mr, c2<double>* __restrict__ Pr,
                                   c2<double>* __restrict__ Pc, c2<float>* __restrict__ Pr32,
                                   c2<float>* __restrict__ Pc32, int64_t n) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const c2<double> m = Mr[v];
        const double d = m.x * m.x + m.y * m.y;
        const c2<double> r = {m.x / d, -m.y / d};
        const c2<double> c = {r.x - 1.0, r.y};
        Pr[v] = r;
        Pc[v] = c;
        if (Pr32) {
            Pr32[v] = c2<float>{(float)r.x, (float)r.y};
            Pc32[v] = c2<float>{(float)c.x, (float)c.y};
        }
    }
}

// The same material for a handle without FP64 data: derived from the complex64
// Mr, computed in FP64 and rounded once.
__global__ void k_precond_material32(const c2<float>* __restrict__ Mr, c2<float>* __restrict__ Pr32,
                                     c2<float>* __restrict__ Pc32, int64_t n) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const double mx = Mr[v].x, my = Mr[v].y;
        const double d = mx * mx + my * my;
        const double rx = mx / d, ry = -my / d;
        Pr32[v] = c2<float>{(float)rx, (float)ry};
        Pc32[v] = c2<float>{(float)(rx - 1.0), (float)ry};
    }
}

// y = Gram * x: the preconditioner when every Mr is 1 (then 1./Mr - 1 = 0 and
// mv_AN reduces to Gram*1*x - 0*(N x), the same value without the FFTs)
template <typename T>
__global__ void k_gram_scale(const c2<T>* __restrict__ x, c2<T>* __restrict__ y, T gram, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = cscale(x[e], gram);
}

static int ensure_precond(Op* h, hipStream_t s) {
    if (h->have_P) return OK;
    const size_t n = (size_t)h->nvox();
    auto A = [&](void** p, size_t bytes) -> int {
        if (*p) return OK;
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            set_error("out of device memory for the preconditioner material");
            return ERR_ALLOC;
        }
        return OK;
    };
    // the FP64 material exists only on handles created with MARIE_PREC_F64
    const bool f64 = h->precisions & PREC_F64;
    const bool f32 = h->precisions & PREC_F32;
    int st = OK;
    if (f64) st = A(&h->Pr64, n * sizeof(c2<double>));
    if (!st && f64) st = A(&h->Pc64, n * sizeof(c2<double>));
    if (!st && f32) st = A(&h->Pr32, n * sizeof(c2<float>));
    if (!st && f32) st = A(&h->Pc32, n * sizeof(c2<float>));
    if (st) return st;
    int64_t g = ((int64_t)n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    if (f64)
        hipLaunchKernelGGL(k_precond_material, dim3((unsigned)g), dim3(256), 0, s, (const c2<double>*)h->Mr64,
                           (c2<double>*)h->Pr64, (c2<double>*)h->Pc64, f32 ? (c2<float>*)h->Pr32 : nullptr,
                           f32 ? (c2<float>*)h->Pc32 : nullptr, (int64_t)n);
    else
        hipLaunchKernelGGL(k_precond_material32, dim3((unsigned)g), dim3(256), 0, s, (const c2<float>*)h->Mr32,
                           (c2<float>*)h->Pr32, (c2<float>*)h->Pc32, (int64_t)n);
    MARIE_HIP(hipGetLastError());
    h->have_P = true;
    return OK;
}

// One product on work buffers Abuf ([3 nrhs][N][M][Px]) and Bbuf
// ([3 nrhs][N][Py][Px]); every element a pass reads was written by an earlier
// pass of the same product, so the buffers need no clearing. `timed`: record
// the per-kernel events when the handle's timing is on. `oscale` multiplies
// the result (folded into the epilogue's two coefficients, no extra pass).
template <typename T>
static int apply_t(Op* h, int op, int transp, const c2<T>* x, c2<T>* y, int nrhs, hipStream_t s, void* Abuf,
                   void* Bbuf, bool timed, double oscale = 1.0) {
    const bool f64 = sizeof(T) == 8;
    const bool kop = (op == OP_K);
    const c2<T>* fN = reinterpret_cast<const c2<T>*>(kop ? (f64 ? h->fK64 : h->fK32) : (f64 ? h->fN64 : h->fN32));
    const bool oct = kop ? h->fK_oct : h->fN_oct;
    const int prod = kop ? 1 : 0;
    const bool pop = (op == OP_P);
    const c2<T>* Mr = reinterpret_cast<const c2<T>*>(pop ? (f64 ? h->Pr64 : h->Pr32) : (f64 ? h->Mr64 : h->Mr32));
    const c2<T>* Mc = reinterpret_cast<const c2<T>*>(pop ? (f64 ? h->Pc64 : h->Pc32) : (f64 ? h->Mc64 : h->Mc32));
    if (pop) op = OP_AN;       // the same product with the preconditioner's material
    const c2<T>* tw = reinterpret_cast<const c2<T>*>(f64 ? h->tw64 : h->tw32);
    const c2<T>* twx = tw;
    const c2<T>* twy = tw + h->Px;
    const c2<T>* twz = tw + h->Px + h->Py;
    c2<T>* A = reinterpret_cast<c2<T>*>(Abuf);
    c2<T>* B = reinterpret_cast<c2<T>*>(Bbuf);
    const int L = h->L, M = h->M, N = h->N, Px = h->Px, Py = h->Py, Pz = h->Pz;
    const int ncomp = 3 * nrhs;
    const int* em = f64 ? h->emax64 : h->emax32;
    const int dxy = f64 ? 16 : 32, dz = f64 ? 8 : 16;
    const int txmax = Variants<T>::txmax;
    // per pass: cfg > 0 -> LDS-stage engine with cfg elements per thread;
    //           cfg < 0 -> register engine with -cfg points per thread (2^k sides)
    bool reg[5];
    const int Pp[5] = {Px, Py, Pz, Py, Px};
    for (int k = 0; k < 5; ++k)
        reg[k] = em[k] < 0 && (k == 2   ? reg_z_supported<T>(Pp[k], -em[k])
                               : (k & 1) ? reg_y_supported<T>(Pp[k], -em[k])
                                         : reg_supported<T>(Pp[k], -em[k]));
    // complex64 column pairs need an even x side (a pair never straddles two
    // rows); the FP64 two-line tiles take any side
    for (int k = 1; k < 5; k += 2)
        if (reg[k] && -em[k] > 200 && (Px & 1) && !f64) reg[k] = false;
    // register-engine points per thread of each pass; a pass whose
    // mixed-radix side (96, 112, 160, 192, 224) has no plan for the requested
    // count takes that side's own (12, 14, 20 points; z: 12 and 10 where its
    // three components fit in registers, at 224 one thread group per
    // component with 14 points, code 414)
    int rc[5];
    for (int k = 0; k < 5; ++k) {
        rc[k] = -em[k];
        if (em[k] < 0 && !reg[k]) {
            const int mix = (Pp[k] == 96 || Pp[k] == 192) ? 12
                          : Pp[k] == 224 ? (k == 2 ? 414 : 14) : Pp[k] == 112 ? 14
                          : Pp[k] == 160 ? (k == 2 ? 10 : 20) : 0;
            if (mix && (k == 2 ? reg_z_supported<T>(Pp[k], mix) : reg_supported<T>(Pp[k], mix))) {
                rc[k] = mix;
                reg[k] = true;
            }
        }
    }
    // the register-engine z pass addresses one component through a 32-bit
    // buffer offset: larger grids take the LDS-stage engine
    if ((int64_t)Px * Py * N * (int64_t)sizeof(c2<T>) > (int64_t)0x7fffffff) reg[2] = false;
    const int dflt[5] = {dxy, dxy, dz, dxy, dxy};
    // LDS-stage variants for a register-engine code whose side the register
    // engine does not cover (the tuner's usual choice on 2-3-7 sides)
    const int fb64[5] = {8, 8, 4, 8, 4}, fb32[5] = {8, 8, 8, 16, 8};
    const AxisPlan* apl[5] = {&h->plan[0], &h->plan[1], &h->plan[2], &h->plan[1], &h->plan[0]};
    int e[5];
    for (int k = 0; k < 5; ++k) {
        int want = em[k] > 0 ? em[k] : (f64 ? fb64[k] : fb32[k]);
        if (em[k] <= 0 && Pp[k] > want * NT) want = dflt[k];
        e[k] = dispatched_emax(f64, k == 2, effective_emax(want, *apl[k], dflt[k]), plan_rset(*apl[k]));
        if (!reg[k] && Pp[k] > e[k] * NT) {
            set_error("circulant side exceeds the single-tile FFT limit of the selected tile variant");
            return ERR_SIZE;
        }
    }
    const int64_t nline = (int64_t)M * N;
    const bool tm = timed && h->timing;
    hipEvent_t* tev = h->ring[h->ring_used];
    if (tm) {
        // timing-only events: no system-scope fence (cache writeback and
        // invalidation) between the kernels they bracket
        for (int k = 0; k < 6; ++k)
            if (!tev[k]) MARIE_HIP(hipEventCreateWithFlags(&tev[k], hipEventDisableSystemFence));
        MARIE_HIP(hipEventRecord(tev[0], s));
    }
    const int conj_in = (op == OP_AN && transp) ? 1 : 0;
    // The x and y passes of several right-hand sides run in groups of columns
    // whose A slabs fit the last-level cache together (256 MB): the y pass
    // then reads what the x pass wrote from that cache, and the inverse x pass
    // what the inverse y pass wrote. The z pass takes every column at once
    // (each fN entry is read once for all of them).
    const int64_t slabA = nline * Px, slabB = (int64_t)N * Py * Px;
    int grp = h->xy_group;
    if (grp <= 0) grp = (int)std::max<int64_t>(1, kXYGroupBytes / (3 * slabA * (int64_t)sizeof(c2<T>)));
    grp = std::min(grp, nrhs);

    auto fwd_xy = [&](int q0, int gq) -> int {
        const int nc = 3 * gq;
        const c2<T>* xq = x + (int64_t)q0 * h->nidx;
        c2<T>* Aq = A + (int64_t)3 * q0 * slabA;
        c2<T>* Bq = B + (int64_t)3 * q0 * slabB;
        // forward x
        if (reg[0]) {
            reg_fwd_x<T>(Px, rc[0], s, xq, Aq, h->pos, Mc, twx, L, M, N, h->nidx, conj_in, nc);
        } else {
            const int lg = pick_lg(Px, e[0] * NT, 16);
            const dim3 g((unsigned)((nline + (1 << lg) - 1) >> lg), nc);
            const size_t sm = (((size_t)Px << lg) + Px) * sizeof(c2<T>);    // tile + twiddles
            MARIE_RSET(h->plan[0], MARIE_EMAX_XY(T, e[0], RS, hipLaunchKernelGGL((k_fwd_x<T, EM, RS>), g, dim3(NT),
                       sm, s, xq, Aq, h->pos, Mc, twx, h->plan[0], L, M, N, lg, h->nidx, conj_in)));
        }
        MARIE_HIP(hipGetLastError());
        if (tm && grp == nrhs) MARIE_HIP(hipEventRecord(tev[1], s));
        // forward y: A[c][k][j<M][i] -> B[c][k][j<Py][i]
        if (reg[1]) {
            reg_strided<T>(false, Py, rc[1], s, Aq, Bq, twy, Px, (int64_t)Px, (int64_t)M * Px, (int64_t)Px,
                           (int64_t)Py * Px, M, Py, slabA, slabB, N, nc);
        } else {
            const int lg = pick_lg(Py, e[1] * NT, txmax);
            const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg), N, nc);
            const size_t sm = (((size_t)Py << lg) + Py) * sizeof(c2<T>);    // tile + twiddles
            MARIE_RSET(h->plan[1], MARIE_EMAX_XY(T, e[1], RS, hipLaunchKernelGGL((k_strided<T, EM, false, RS>), g,
                       dim3(NT), sm, s, Aq, Bq, twy, h->plan[1], Px, (int64_t)Px, (int64_t)M * Px, (int64_t)Px,
                       (int64_t)Py * Px, M, Py, lg, slabA, slabB)));
        }
        MARIE_HIP(hipGetLastError());
        return OK;
    };
    auto inv_yx = [&](int q0, int gq) -> int {
        const int nc = 3 * gq;
        const c2<T>* xq = x + (int64_t)q0 * h->nidx;
        c2<T>* yq = y + (int64_t)q0 * h->nidx;
        c2<T>* Aq = A + (int64_t)3 * q0 * slabA;
        const c2<T>* Bq = B + (int64_t)3 * q0 * slabB;
        // inverse y: B[c][k][j<Py][i] -> A[c][k][j<M][i]
        if (reg[3]) {
            reg_strided<T>(true, Py, rc[3], s, Bq, Aq, twy, Px, (int64_t)Px, (int64_t)Py * Px, (int64_t)Px,
                           (int64_t)M * Px, Py, M, slabB, slabA, N, nc);
        } else {
            const int lg = pick_lg(Py, e[3] * NT, txmax);
            const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg), N, nc);
            const size_t sm = (((size_t)Py << lg) + Py) * sizeof(c2<T>);    // tile + twiddles
            MARIE_RSET(h->plan[1], MARIE_EMAX_XY(T, e[3], RS, hipLaunchKernelGGL((k_strided<T, EM, true, RS>), g,
                       dim3(NT), sm, s, Bq, Aq, twy, h->plan[1], Px, (int64_t)Px, (int64_t)Py * Px, (int64_t)Px,
                       (int64_t)M * Px, Py, M, lg, slabB, slabA)));
        }
        MARIE_HIP(hipGetLastError());
        if (tm && grp == nrhs) MARIE_HIP(hipEventRecord(tev[4], s));
        // inverse x + crop + epilogue
        const T scale = T(oscale / ((double)Px * (double)Py * (double)Pz));
        const T gram = T(h->Gram * oscale);
        if (reg[4]) {
            reg_inv_x<T>(Px, rc[4], s, Aq, yq, xq, h->pos, Mr, Mc, twx, L, M, N, h->nidx, op, transp, gram, scale,
                         nc);
        } else {
            const int lg = pick_lg(Px, e[4] * NT, 16);
            const dim3 g((unsigned)((nline + (1 << lg) - 1) >> lg), nc);
            const size_t sm = (((size_t)Px << lg) + Px) * sizeof(c2<T>);    // tile + twiddles
            MARIE_RSET(h->plan[0], MARIE_EMAX_XY(T, e[4], RS, hipLaunchKernelGGL((k_inv_x<T, EM, RS>), g, dim3(NT),
                       sm, s, Aq, yq, xq, h->pos, Mr, Mc, twx, h->plan[0], L, M, N, lg, h->nidx, op, transp,
                       gram, scale)));
        }
        MARIE_HIP(hipGetLastError());
        return OK;
    };

    // with grouped columns the timing ring cannot split the interleaved
    // passes: fwd_x reads as both forward passes (fwd_y zero), inv_x as both
    // inverse passes (inv_y zero)
    for (int q0 = 0; q0 < nrhs; q0 += grp) {
        const int st = fwd_xy(q0, std::min(grp, nrhs - q0));
        if (st != OK) return st;
    }
    if (tm && grp < nrhs) MARIE_HIP(hipEventRecord(tev[1], s));
    if (tm) MARIE_HIP(hipEventRecord(tev[2], s));
    // z: fused forward, tensor product, inverse
    if (reg[2]) {
        reg_z<T>(oct, prod, Pz, rc[2], s, B, fN, twz, Px, Py, N, nrhs);
    } else {
        const int lg = pick_lg(Pz, e[2] * NT, txmax);
        const dim3 g((unsigned)((Px + (1 << lg) - 1) >> lg) * Py * nrhs);
        const size_t sm = (((size_t)Pz << lg) + Pz) * sizeof(c2<T>);    // tile + twiddles
        if (oct && !kop)
            MARIE_RSET(h->plan[2], MARIE_EMAX_Z(T, e[2], RS, hipLaunchKernelGGL((k_z<T, EM, RS, true, 0>), g,
                       dim3(NT), sm, s, B, fN, twz, h->plan[2], Px, Py, N, lg, nrhs)));
        else if (!kop)
            MARIE_RSET(h->plan[2], MARIE_EMAX_Z(T, e[2], RS, hipLaunchKernelGGL((k_z<T, EM, RS, false, 0>), g,
                       dim3(NT), sm, s, B, fN, twz, h->plan[2], Px, Py, N, lg, nrhs)));
        else if (oct)
            MARIE_RSET(h->plan[2], MARIE_EMAX_Z(T, e[2], RS, hipLaunchKernelGGL((k_z<T, EM, RS, true, 1>), g,
                       dim3(NT), sm, s, B, fN, twz, h->plan[2], Px, Py, N, lg, nrhs)));
        else
            MARIE_RSET(h->plan[2], MARIE_EMAX_Z(T, e[2], RS, hipLaunchKernelGGL((k_z<T, EM, RS, false, 1>), g,
                       dim3(NT), sm, s, B, fN, twz, h->plan[2], Px, Py, N, lg, nrhs)));
    }
    MARIE_HIP(hipGetLastError());
    if (tm) MARIE_HIP(hipEventRecord(tev[3], s));
    if (tm && grp < nrhs) MARIE_HIP(hipEventRecord(tev[4], s));
    for (int q0 = 0; q0 < nrhs; q0 += grp) {
        const int st = inv_yx(q0, std::min(grp, nrhs - q0));
        if (st != OK) return st;
    }
    if (tm) {
        MARIE_HIP(hipEventRecord(tev[5], s));
        if (++h->ring_used == Op::kTimingRing) return timing_flush(h);
    }
    return OK;
}

int timing_flush(Op* h) {
    if (h->ring_used == 0) return OK;
    MARIE_HIP(hipEventSynchronize(h->ring[h->ring_used - 1][5]));
    for (int r = 0; r < h->ring_used; ++r)
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            MARIE_HIP(hipEventElapsedTime(&ms, h->ring[r][k], h->ring[r][k + 1]));
            h->kernel_ms[k] += ms;
        }
    h->kernel_calls += h->ring_used;
    h->ring_used = 0;
    return OK;
}

// state and argument checks shared by both product entry points
static int apply_check(Op* h, int op, int transp, int prec, int nrhs) {
    const bool have_op = (op == OP_K) ? h->have_fK : h->have_fN;
    if (!have_op || !h->have_idx || ((op == OP_AN || op == OP_P) && !h->have_mat)) {
        set_error(op == OP_K ? "K operator or index not set" : "operator, index or material not set");
        return ERR_STATE;
    }
    if (op == OP_K && transp) {
        set_error("the K operator has no transposed product");
        return ERR_ARG;
    }
    if (!(h->precisions & prec)) {
        set_error("requested precision was not enabled at create");
        return ERR_STATE;
    }
    // the z pass runs one workgroup per (column tile, row, rhs) in a 1D grid;
    // keep the grid's work-item count within 32 bits for the smallest tile
    if ((int64_t)h->Px * h->Py * nrhs * NT >= (int64_t)1 << 32) {
        set_error("too many right-hand sides for one product on this grid");
        return ERR_ARG;
    }
    return OK;
}

// y = Gram x on the compact vectors (OP_P when every Mr is 1)
static int gram_scale(Op* h, int prec, const void* x, void* y, int nrhs, hipStream_t s) {
    const int64_t n = h->nidx * nrhs;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_gram_scale<double>, dim3((unsigned)g), dim3(256), 0, s, (const c2<double>*)x,
                           (c2<double>*)y, h->Gram, n);
    else
        hipLaunchKernelGGL(k_gram_scale<float>, dim3((unsigned)g), dim3(256), 0, s, (const c2<float>*)x,
                           (c2<float>*)y, (float)h->Gram, n);
    MARIE_HIP(hipGetLastError());
    return OK;
}

int apply(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s) {
    const int chk = apply_check(h, op, transp, prec, nrhs);
    if (chk) return chk;
    if (op == OP_P) {
        if (h->mr_unit) return gram_scale(h, prec, x, y, nrhs, s);
        const int stp = ensure_precond(h, s);
        if (stp) return stp;
    }
    int st = ensure_work(h, nrhs);
    if (st) return st;
    if (prec == PREC_F64)
        return apply_t<double>(h, op, transp, (const c2<double>*)x, (c2<double>*)y, nrhs, s, h->A, h->B, true);
    return apply_t<float>(h, op, transp, (const c2<float>*)x, (c2<float>*)y, nrhs, s, h->A, h->B, true);
}

int apply_scaled(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s,
                 double oscale) {
    if (op == OP_P) {
        set_error("the scaled product covers MARIE_OP_AN, N, L and K");
        return ERR_ARG;
    }
    const int chk = apply_check(h, op, transp, prec, nrhs);
    if (chk) return chk;
    int st = ensure_work(h, nrhs);
    if (st) return st;
    if (prec == PREC_F64)
        return apply_t<double>(h, op, transp, (const c2<double>*)x, (c2<double>*)y, nrhs, s, h->A, h->B, true,
                               oscale);
    return apply_t<float>(h, op, transp, (const c2<float>*)x, (c2<float>*)y, nrhs, s, h->A, h->B, true, oscale);
}

// ------------------------------------------------- allocation-free product
// Work buffers of marie_vie_apply_device_noalloc: one file-scope device array
// per GPU, held until the process ends. The host name `scratch` is not a
// device address; its address on each device comes from hipGetSymbolAddress
// (scratch_init, called by marie_vie_create with that device current) and
// only that pointer is passed to kernels. Every call on a device shares its
// array, so calls that overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)128 << 20;     // 128 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device ordinal beyond the scratch table");
        return ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return OK;
    void* p = nullptr;
    MARIE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
    return OK;
}

size_t scratch_capacity() { return kScratchBytes; }

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of A and B for one product (offset of B = align256 of A's bytes)
size_t noalloc_bytes(const Op* h, int prec, int nrhs) {
    const size_t es = prec == PREC_F64 ? sizeof(c2<double>) : sizeof(c2<float>);
    const size_t na = (size_t)3 * nrhs * h->Px * h->M * h->N;
    const size_t nb = (size_t)3 * nrhs * h->Px * h->Py * h->N;
    return align256(na * es) + nb * es;
}

int apply_noalloc(Op* h, int op, int transp, int prec, const void* x, void* y, int nrhs, hipStream_t s) {
    if (nrhs >= 1 && noalloc_bytes(h, prec, nrhs) > kScratchBytes) {
        set_error("the product's work buffers (" + std::to_string(noalloc_bytes(h, prec, nrhs)) +
                  " bytes) exceed the allocation-free scratch of " + std::to_string(kScratchBytes) + " bytes");
        return ERR_SIZE;
    }
    const int chk = apply_check(h, op, transp, prec, nrhs);
    if (chk) return chk;
    if (op == OP_P) {
        if (h->mr_unit) return gram_scale(h, prec, x, y, nrhs, s);
        if (!h->have_P) {
            set_error("the preconditioner's material is derived on the first MARIE_OP_P product of "
                      "marie_vie_apply_device (it allocates); run one there before this entry point");
            return ERR_STATE;
        }
    }
    void* base = (h->device >= 0 && h->device < kMaxDevices) ? g_scratch_ptr[h->device] : nullptr;
    if (!base) {
        set_error("allocation-free scratch not initialised on this device");
        return ERR_STATE;
    }
    const size_t es = prec == PREC_F64 ? sizeof(c2<double>) : sizeof(c2<float>);
    void* A = base;
    void* B = static_cast<unsigned char*>(base) + align256((size_t)3 * nrhs * h->Px * h->M * h->N * es);
    if (prec == PREC_F64)
        return apply_t<double>(h, op, transp, (const c2<double>*)x, (c2<double>*)y, nrhs, s, A, B, false);
    return apply_t<float>(h, op, transp, (const c2<float>*)x, (c2<float>*)y, nrhs, s, A, B, false);
}

// Times every compiled tile variant of each pass with HIP events on the
// handle's stream (real buffers, `iters` products each) and picks the fastest
// per pass, then revisits each pass with the others at their picks and keeps
// the code that gives the fastest whole product. Results are left in
// h->emax64 / h->emax32; best_ms5 gets the final configuration's pass times.
int autotune(Op* h, int prec, int iters, double* best_ms5) {
    const bool f64 = prec == PREC_F64;
    int* em = f64 ? h->emax64 : h->emax32;
    const int nxy = 3, nz = 2;
    const int* vxy = f64 ? Variants<double>::xy : Variants<float>::xy;
    const int* vz = f64 ? Variants<double>::z : Variants<float>::z;
    const size_t es = f64 ? sizeof(c2<double>) : sizeof(c2<float>);
    void *x = nullptr, *y = nullptr;
    const size_t n = (size_t)(h->nidx > 0 ? h->nidx : 1);
    if (hipMalloc(&x, n * es) != hipSuccess || hipMalloc(&y, n * es) != hipSuccess ||
        hipMemsetAsync(x, 0, n * es, h->stream) != hipSuccess) {
        (void)hipGetLastError();
        if (x) (void)hipFree(x);
        if (y) (void)hipFree(y);
        set_error("out of device memory for the autotuner's vectors");
        return ERR_ALLOC;
    }
    double best[5] = {1e30, 1e30, 1e30, 1e30, 1e30};
    int pick[5] = {em[0], em[1], em[2], em[3], em[4]};
    const bool was = h->timing;
    int st = OK;
    // candidate configurations: the LDS-stage engine's tile variants, then the
    // register engine's points-per-thread variants (power-of-two sides only)
    int cands[16][5];
    int nc = 0;
    for (int v = 0; v < nxy; ++v) {
        const int exy = vxy[v], ez = vz[v < nz ? v : nz - 1];
        if (h->Px > exy * NT || h->Py > exy * NT || h->Pz > ez * NT) continue;
        const int c[5] = {exy, exy, ez, exy, exy};
        for (int k = 0; k < 5; ++k) cands[nc][k] = c[k];
        ++nc;
    }
    const int regE[3] = {4, 8, 16};
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 5; ++k) cands[nc + v][k] = -regE[v];
    nc += 3;
    // register-engine plans of the mixed-radix sides (96, 192: 12 or 24
    // points per thread; 112, 224: 14 or 28; 160: 10 or 20); a pass whose
    // side has no such plan runs its default in these candidates
    const int mixE[6] = {12, 24, 14, 28, 10, 20};
    const int Pp3[3] = {h->Px, h->Py, h->Pz};
    for (int v = 0; v < 6 && nc < 16; ++v) {
        bool any = false;
        for (int a = 0; a < 3; ++a)
            any = any || (f64 ? reg_supported<double>(Pp3[a], mixE[v]) : reg_supported<float>(Pp3[a], mixE[v]));
        if (!any) continue;
        for (int k = 0; k < 5; ++k) cands[nc][k] = -mixE[v];
        ++nc;
    }
    // y passes on two lines per thread (code 200 + E): complex64 column pairs
    // with x and z passes at 16 points; FP64 columns TX apart at 8 points
    if (!f64 && reg_y_supported<float>(h->Py, 216) && nc < 16) {
        for (int k = 0; k < 5; ++k) cands[nc][k] = -16;
        cands[nc][1] = cands[nc][3] = -216;
        ++nc;
    }
    if (f64 && reg_y_supported<double>(h->Py, 208) && nc < 16) {
        for (int k = 0; k < 5; ++k) cands[nc][k] = -8;
        cands[nc][1] = cands[nc][3] = -208;
        ++nc;
    }
    // z pass with one thread group per component (code 100 * TX + E)
    const int zsplit[5] = {408, 404, 416, 808, 414};
    for (int v = 0; v < 5 && nc < 16; ++v) {
        if (!reg_z_supported<double>(h->Pz, zsplit[v]) && f64) continue;
        if (!f64 && !reg_z_supported<float>(h->Pz, zsplit[v])) continue;
        for (int k = 0; k < 5; ++k) cands[nc][k] = -8;
        cands[nc][2] = -zsplit[v];
        ++nc;
    }
    const bool verbose = getenv("MARIE_TUNE_VERBOSE") != nullptr;
    for (int ci = 0; ci < nc && st == OK; ++ci) {
        const int cj = ci;
        for (int k = 0; k < 5; ++k) em[k] = cands[cj][k];
        h->timing = false;
        st = apply(h, OP_AN, 0, prec, x, y, 1, h->stream);      // warm-up
        if (st) break;
        h->timing = true;
        for (int k = 0; k < 8; ++k) h->kernel_ms[k] = 0.0;
        h->kernel_calls = 0;
        for (int it = 0; it < iters && st == OK; ++it) st = apply(h, OP_AN, 0, prec, x, y, 1, h->stream);
        if (st == OK) st = timing_flush(h);
        for (int k = 0; k < 5 && st == OK; ++k) {
            const double ms = h->kernel_ms[k] / (h->kernel_calls ? h->kernel_calls : 1);
            if (ms < best[k]) {
                best[k] = ms;
                pick[k] = cands[cj][k];
            }
        }
        if (verbose)
            fprintf(stderr, "[marie tune] prec %d cfg [%d %d %d %d %d] : %.4f %.4f %.4f %.4f %.4f ms\n",
                    prec, em[0], em[1], em[2], em[3], em[4], h->kernel_ms[0] / iters, h->kernel_ms[1] / iters,
                    h->kernel_ms[2] / iters, h->kernel_ms[3] / iters, h->kernel_ms[4] / iters);
    }
    // refinement: a pass's time depends on what its neighbours left in the
    // caches (after the slow one-group-per-component z candidates the y
    // passes measure faster than beside the z pass they will run with), so
    // each pass is revisited with the others at their picks, every code that
    // pass had among the candidates, and the whole product's time decides
    auto product_ms = [&](double* per) -> int {
        h->timing = false;
        int s2 = apply(h, OP_AN, 0, prec, x, y, 1, h->stream);
        if (s2) return s2;
        h->timing = true;
        for (int k = 0; k < 8; ++k) h->kernel_ms[k] = 0.0;
        h->kernel_calls = 0;
        for (int it = 0; it < iters && s2 == OK; ++it) s2 = apply(h, OP_AN, 0, prec, x, y, 1, h->stream);
        if (s2 == OK) s2 = timing_flush(h);
        for (int k = 0; k < 5; ++k) per[k] = h->kernel_ms[k] / (h->kernel_calls ? h->kernel_calls : 1);
        return s2;
    };
    if (st == OK) {
        double per[5];
        for (int k = 0; k < 5; ++k) em[k] = pick[k];
        st = product_ms(per);
        double total = 0.0;
        for (int k = 0; k < 5; ++k) {
            total += per[k];
            best[k] = per[k];
        }
        for (int k = 0; k < 5 && st == OK; ++k) {
            int seen[16], ns = 0;
            for (int ci = 0; ci < nc; ++ci) {
                bool dup = cands[ci][k] == pick[k];
                for (int q = 0; q < ns; ++q) dup = dup || seen[q] == cands[ci][k];
                if (!dup) seen[ns++] = cands[ci][k];
            }
            for (int q = 0; q < ns && st == OK; ++q) {
                for (int k2 = 0; k2 < 5; ++k2) em[k2] = pick[k2];
                em[k] = seen[q];
                double p2[5];
                st = product_ms(p2);
                if (st) break;
                double t2 = 0.0;
                for (int k2 = 0; k2 < 5; ++k2) t2 += p2[k2];
                if (verbose)
                    fprintf(stderr, "[marie tune] refine pass %d code %d: %.4f ms (best %.4f)\n", k, seen[q], t2,
                            total);
                if (t2 < total) {
                    total = t2;
                    pick[k] = seen[q];
                    for (int k2 = 0; k2 < 5; ++k2) best[k2] = p2[k2];
                }
            }
        }
    }
    for (int k = 0; k < 5; ++k) em[k] = pick[k];
    h->timing = was;
    for (int k = 0; k < 8; ++k) h->kernel_ms[k] = 0.0;
    h->kernel_calls = 0;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(x);
    (void)hipFree(y);
    if (best_ms5)
        for (int k = 0; k < 5; ++k) best_ms5[k] = best[k];
    return st;
}

}  // namespace marie
