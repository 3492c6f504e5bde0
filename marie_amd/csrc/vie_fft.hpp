// LDS-staged mixed-radix Stockham FFT engine for gfx950 (wave64).
//
// A workgroup owns a tile of TX lines of length P held in LDS with layout
// element(p, b) = buf[p * TX + b]: the line index b is the fast index, so the
// global<->LDS copies of strided axes are coalesced over b, and every stage's
// LDS reads and writes are consecutive over b (no bank conflicts for 8- or
// 16-byte elements). Each stage reads all its butterfly inputs into registers,
// synchronises, then computes and writes the outputs in place; twiddles come
// from a per-axis table tw[k] = exp(-2 pi i k / P) staged in LDS.
//
// Supported radices: 2, 3, 4, 5, 8 with closed-form butterflies and any odd
// prime up to 13 through a table-driven butterfly. Sizes are plans of these.
#pragma once
#include <hip/hip_runtime.h>

namespace marie {

// Workgroup barrier for the LDS exchanges: waits for this wave's LDS
// accesses (lgkmcnt) only, so global loads already issued (the next field
// component) stay in flight across it. The kernels that use it communicate
// between threads through LDS only. The empty asm statements keep the
// compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);      // vmcnt, expcnt at their maximum; lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <typename T> struct __align__(2 * sizeof(T)) c2 {
    T x, y;
};

template <typename T> __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T> __device__ __forceinline__ c2<T> cmulc(c2<T> a, c2<T> b) {
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T> __device__ __forceinline__ c2<T> cscale(c2<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T> __device__ __forceinline__ c2<T> cconj(c2<T> a) { return {a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ c2<T> mul_mi(c2<T> a) {
    return INV ? c2<T>{-a.y, a.x} : c2<T>{a.y, -a.x};
}

constexpr int kMaxStages = 16;
constexpr int kMaxRadix = 13;

struct AxisPlan {
    int P;            // line length
    int nst;          // number of stages
    int radix[kMaxStages];
};

// ---------------------------------------------------------------- butterflies
template <typename T, bool INV> __device__ __forceinline__ void dft2(c2<T>* a) {
    c2<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

template <typename T, bool INV> __device__ __forceinline__ void dft4(c2<T>* a) {
    c2<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    c2<T> t2 = cadd(a[1], a[3]), t3 = mul_mi<T, INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft8(c2<T>* a) {
    const T h = T(0.70710678118654752440084436210484903928);
    c2<T> e[4] = {a[0], a[2], a[4], a[6]};
    c2<T> o[4] = {a[1], a[3], a[5], a[7]};
    dft4<T, INV>(e);
    dft4<T, INV>(o);
    // o[k] *= w8^k, w8 = exp(-+ i pi/4)
    c2<T> o1 = INV ? c2<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                   : c2<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
    c2<T> o2 = mul_mi<T, INV>(o[2]);
    c2<T> o3 = INV ? c2<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                   : c2<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
}

template <typename T, bool INV> __device__ __forceinline__ void dft3(c2<T>* a) {
    const T s3 = T(0.86602540378443864676372317075293618347);   // sin(2pi/3)
    c2<T> t = cadd(a[1], a[2]);
    c2<T> d = csub(a[1], a[2]);
    c2<T> m = {a[0].x - T(0.5) * t.x, a[0].y - T(0.5) * t.y};
    c2<T> r = mul_mi<T, INV>(cscale(d, s3));
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
}

template <typename T, bool INV> __device__ __forceinline__ void dft5(c2<T>* a) {
    const T c1 = T(0.30901699437494742410229341718281905886);    // cos(2pi/5)
    const T c2_ = T(-0.80901699437494742410229341718281905886);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938214340);    // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907276860);    // sin(4pi/5)
    c2<T> t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    c2<T> t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
    c2<T> b1 = {a[0].x + c1 * t1.x + c2_ * t2.x, a[0].y + c1 * t1.y + c2_ * t2.y};
    c2<T> b2 = {a[0].x + c2_ * t1.x + c1 * t2.x, a[0].y + c2_ * t1.y + c1 * t2.y};
    c2<T> u1 = mul_mi<T, INV>(c2<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
    c2<T> u2 = mul_mi<T, INV>(c2<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(b1, u1);
    a[4] = csub(b1, u1);
    a[2] = cadd(b2, u2);
    a[3] = csub(b2, u2);
}

// Odd prime radix R (7, 11, 13): X_k = a0 + sum_n [cos(2 pi n k/R)(a_n + a_{R-n})
//   -+ i sin(2 pi n k / R)(a_n - a_{R-n})], twiddles from the axis table.
template <typename T, bool INV, int R>
__device__ __forceinline__ void dft_odd(c2<T>* a, const c2<T>* tw, int P) {
    constexpr int H = (R - 1) / 2;
    c2<T> sp[H], sm[H];
    c2<T> x0 = a[0];
    c2<T> s0 = a[0];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        sp[n - 1] = cadd(a[n], a[R - n]);
        sm[n - 1] = csub(a[n], a[R - n]);
        s0 = cadd(s0, sp[n - 1]);
    }
    const int step = P / R;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        c2<T> re = x0, im = {T(0), T(0)};
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            c2<T> w = tw[((n * k) % R) * step];   // (cos, -sin) of 2 pi nk/R
            re.x += w.x * sp[n - 1].x;
            re.y += w.x * sp[n - 1].y;
            im.x -= w.y * sm[n - 1].x;             // sin * (a_n - a_{R-n})
            im.y -= w.y * sm[n - 1].y;
        }
        c2<T> u = mul_mi<T, INV>(im);
        a[k] = cadd(re, u);
        a[R - k] = csub(re, u);
    }
    a[0] = s0;
}

template <typename T, bool INV, int R>
__device__ __forceinline__ void dft(c2<T>* a, const c2<T>* tw, int P) {
    if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 3) dft3<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 5) dft5<T, INV>(a);
    else if constexpr (R == 8) dft8<T, INV>(a);
    else dft_odd<T, INV, R>(a, tw, P);
}

// ---------------------------------------------------------------- one stage
// Stockham radix-R stage on a tile of `ntx` (=1<<lg) lines, sub-transform size Ns.
// Input of butterfly jj of line b: x[jj + r P/R]; output: y[(jj/Ns)Ns R + jj%Ns + r Ns].
template <typename T, int R, int EMAX, bool INV>
__device__ __forceinline__ void stage(c2<T>* buf, const c2<T>* tw, int P, int Ns, int lg) {
    constexpr int QMAX = (EMAX + R - 1) / R;
    const int nb = P / R;
    const int total = nb << lg;
    const int msk = (1 << lg) - 1;
    c2<T> v[QMAX][R];
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int u = threadIdx.x + q * blockDim.x;
        if (u < total) {
            const int b = u & msk, jj = u >> lg;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[((jj + r * nb) << lg) + b];
        }
    }
    lds_barrier();
    const int tstride = P / (Ns * R);
#pragma unroll
    for (int q = 0; q < QMAX; ++q) {
        const int u = threadIdx.x + q * blockDim.x;
        if (u < total) {
            const int b = u & msk, jj = u >> lg;
            const int k = jj % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    c2<T> w = tw[r * k * tstride];
                    v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
                }
            }
            dft<T, INV, R>(v[q], tw, P);
            const int jo = (jj - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[((jo + r * Ns) << lg) + b] = v[q][r];
        }
    }
    lds_barrier();
}

// Radix sets: a kernel is instantiated for the smallest set its plan needs,
// which keeps the register allocation of the common power-of-two case small.
// (7 apart from 11 and 13: the largest radix sets the kernel's register
// count, and 7-smooth sides are the common case)
constexpr int RSET_POW2 = 0;   // 2, 4, 8
constexpr int RSET_SMALL = 1;  // + 3, 5
constexpr int RSET_SEVEN = 2;  // + 7
constexpr int RSET_ALL = 3;    // + 11, 13

__host__ __device__ inline int plan_rset(const AxisPlan& pl) {
    int s = RSET_POW2;
    for (int i = 0; i < pl.nst; ++i) {
        const int R = pl.radix[i];
        if (R == 3 || R == 5) s = s < RSET_SMALL ? RSET_SMALL : s;
        if (R == 7) s = s < RSET_SEVEN ? RSET_SEVEN : s;
        if (R == 11 || R == 13) s = RSET_ALL;
    }
    return s;
}

// Full transform of the tile in LDS. buf and tw live in LDS or global; the
// caller synchronises before (data loaded) and this returns synchronised.
// The barriers order LDS accesses only (lds_barrier): callers communicate
// between threads through LDS alone, so their global loads stay in flight.
template <typename T, int EMAX, bool INV, int RSET>
__device__ void tile_fft(c2<T>* buf, const c2<T>* tw, const AxisPlan& pl, int lg) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.radix[s];
        if (R == 2) stage<T, 2, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        else if (R == 4) stage<T, 4, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        else if (R == 8) stage<T, 8, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        if constexpr (RSET >= RSET_SMALL) {
            if (R == 3) stage<T, 3, EMAX, INV>(buf, tw, pl.P, Ns, lg);
            else if (R == 5) stage<T, 5, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        }
        if constexpr (RSET >= RSET_SEVEN) {
            if (R == 7) stage<T, 7, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        }
        if constexpr (RSET >= RSET_ALL) {
            if (R == 11) stage<T, 11, EMAX, INV>(buf, tw, pl.P, Ns, lg);
            else if (R == 13) stage<T, 13, EMAX, INV>(buf, tw, pl.P, Ns, lg);
        }
        Ns *= R;
    }
}

// Calls STMT with the compile-time constant RS set to the plan's radix set.
#define MARIE_RSET(plan, STMT)                                             \
    do {                                                                   \
        switch (::marie::plan_rset(plan)) {                                \
            case ::marie::RSET_POW2: { constexpr int RS = 0; STMT; } break; \
            case ::marie::RSET_SMALL: { constexpr int RS = 1; STMT; } break;\
            case ::marie::RSET_SEVEN: { constexpr int RS = 2; STMT; } break;\
            default: { constexpr int RS = 3; STMT; } break;                 \
        }                                                                  \
    } while (0)

// The product of the z pass at one frequency, in place on the three transformed
// components x0, x1, x2. f points at the operator entry of component 0 (the
// others follow at stride fs); OCT: the entry was read from the octant of a
// circulant with per-axis parity and fx, fy, fz say which axes were mirrored.
//   PROD 0, the N operator: symmetric 3x3 tensor of the 6 components
//     (xx, xy, xz, yy, yz, zz; mv_AN.m:108-123), xy odd under x and y
//     reflection, xz under x and z, yz under y and z;
//   PROD 1, the K operator: y = [Kz x1 - Ky x2, -Kz x0 + Kx x2, Ky x0 - Kx x1]
//     (mv_K.m:66-80), component c odd under reflection of axis c.
template <typename T, int PROD, bool OCT>
__device__ __forceinline__ void freq_product(const c2<T>* __restrict__ f, int64_t fs, bool fx, bool fy, bool fz,
                                             c2<T>& x0, c2<T>& x1, c2<T>& x2) {
    if constexpr (PROD == 0) {
        c2<T> f0 = f[0], f1 = f[fs], f2 = f[2 * fs], f3 = f[3 * fs], f4 = f[4 * fs], f5 = f[5 * fs];
        if (OCT) {
            if (fx != fy) f1 = c2<T>{-f1.x, -f1.y};
            if (fx != fz) f2 = c2<T>{-f2.x, -f2.y};
            if (fy != fz) f4 = c2<T>{-f4.x, -f4.y};
        }
        const c2<T> a = x0, b = x1, c = x2;
        x0 = cadd(cadd(cmul(f0, a), cmul(f1, b)), cmul(f2, c));
        x1 = cadd(cadd(cmul(f1, a), cmul(f3, b)), cmul(f4, c));
        x2 = cadd(cadd(cmul(f2, a), cmul(f4, b)), cmul(f5, c));
    } else {
        c2<T> k0 = f[0], k1 = f[fs], k2 = f[2 * fs];
        if (OCT) {
            if (fx) k0 = c2<T>{-k0.x, -k0.y};
            if (fy) k1 = c2<T>{-k1.x, -k1.y};
            if (fz) k2 = c2<T>{-k2.x, -k2.y};
        }
        const c2<T> a = x0, b = x1, c = x2;
        x0 = csub(cmul(k2, b), cmul(k1, c));
        x1 = csub(cmul(k0, c), cmul(k2, a));
        x2 = csub(cmul(k1, a), cmul(k0, b));
    }
}

// Row `a` of freq_product: y_a from x0, x1, x2 (same operations, same order,
// for the thread group that owns component a).
template <typename T, int PROD, bool OCT>
__device__ __forceinline__ c2<T> freq_row(const c2<T>* __restrict__ f, int64_t fs, bool fx, bool fy, bool fz, int a,
                                          c2<T> x0, c2<T> x1, c2<T> x2) {
    if constexpr (PROD == 0) {
        // rows {0,1,2}, {1,3,4}, {2,4,5} of the symmetric tensor
        const int i0 = a == 0 ? 0 : (a == 1 ? 1 : 2);
        const int i1 = a == 0 ? 1 : (a == 1 ? 3 : 4);
        const int i2 = a == 0 ? 2 : (a == 1 ? 4 : 5);
        c2<T> g0 = f[i0 * fs], g1 = f[i1 * fs], g2 = f[i2 * fs];
        if (OCT) {
            // component 1 (xy) flips when fx != fy, 2 (xz) when fx != fz, 4 (yz) when fy != fz
            const bool s1 = fx != fy, s2 = fx != fz, s4 = fy != fz;
            if ((i0 == 1 && s1) || (i0 == 2 && s2)) g0 = c2<T>{-g0.x, -g0.y};
            if ((i1 == 1 && s1) || (i1 == 4 && s4)) g1 = c2<T>{-g1.x, -g1.y};
            if ((i2 == 2 && s2) || (i2 == 4 && s4)) g2 = c2<T>{-g2.x, -g2.y};
        }
        return cadd(cadd(cmul(g0, x0), cmul(g1, x1)), cmul(g2, x2));
    } else {
        // y0 = k2 x1 - k1 x2, y1 = k0 x2 - k2 x0, y2 = k1 x0 - k0 x1
        const int ip = a == 0 ? 2 : (a == 1 ? 0 : 1);      // multiplies the next component
        const int im = a == 0 ? 1 : (a == 1 ? 2 : 0);      // multiplies the one after
        c2<T> kp = f[ip * fs], km = f[im * fs];
        if (OCT) {
            const bool fl[3] = {fx, fy, fz};
            if (fl[ip]) kp = c2<T>{-kp.x, -kp.y};
            if (fl[im]) km = c2<T>{-km.x, -km.y};
        }
        const c2<T> xn = a == 0 ? x1 : (a == 1 ? x2 : x0);
        const c2<T> xa = a == 0 ? x2 : (a == 1 ? x0 : x1);
        return csub(cmul(kp, xn), cmul(km, xa));
    }
}

// Dispatch order pairing mirror images: 0, 1, n-1, 2, n-2, ..., n/2 (n even).
__device__ __forceinline__ int mirror_order2(int y, int n) {
    if (y == 0 || (n & 1)) return y;
    if (y == n - 1) return n / 2;
    const int p = (y + 1) >> 1;
    return (y & 1) ? p : n - p;
}

// Tile of the z pass for workgroup `lid` of a 1D grid of gx * Py * nrhs tiles
// (gx column blocks, Py rows, nrhs right-hand sides).
//
// The 8 XCDs of the chip each have their own L2, and workgroups are handed to
// them round-robin (lid % 8). Consecutive logical tiles are given to one XCD:
// XCD x runs the x-th contiguous range of tiles, in order. Within that order
// the right-hand sides of a tile come first, then the column block and its
// mirror (bx, gx-1-bx), then rows in the order 0, 1, Py-1, 2, Py-2, ...; so
// the (up to 8 nrhs) tiles that read the same octant entries of fN run back
// to back on one XCD and share its L2. Any lid -> tile map here is a
// bijection, so the result does not depend on how the hardware dispatches.
__device__ __forceinline__ void zpass_tile(int lid, int nt, int gx, int Py, int nrhs, int& bx, int& j, int& rhs) {
    const int xcd = lid & 7, s = lid >> 3;
    const int q = nt >> 3, r = nt & 7;
    const int lam = xcd * q + (xcd < r ? xcd : r) + s;
    rhs = lam % nrhs;
    const int u = lam / nrhs;
    const int v = u % gx;
    bx = (v & 1) ? gx - 1 - (v >> 1) : (v >> 1);
    j = mirror_order2(u / gx, Py);
}

}  // namespace marie
