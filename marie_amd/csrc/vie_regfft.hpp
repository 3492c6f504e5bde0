// Register-resident Stockham FFT for power-of-two lines (gfx950, wave64).
//
// A line of P points is owned by Tl = P/E threads; thread t holds E points in
// registers. Stage s (radix R | E) gives thread t the butterflies
// jj = t + Tl*q, q < E/R; inputs x[jj + r P/R] sit in slot v[q*R + r].
// Between stages the outputs are exchanged through LDS (one write, barrier,
// one read); the butterflies, twiddles and the first load / last store stay in
// registers. Because the outputs of a last stage with radix R land in exactly
// the slot layout a first stage with radix R reads, a forward transform with
// radices (R1..Rk) hands its result to a pointwise operation and then to an
// inverse transform with radices (Rk..R1) without touching LDS: slot
// v[q*R + r] holds position t + Tl*(q + (E/R) r) for the first input and the
// last output layouts alike.
#pragma once
#include "vie_fft.hpp"

namespace marie {

struct RegPlan {
    int P, E, Tl;
    int nst;
    int radix[12];
};

// e^{-+ i pi m / 8} for the radix-16 butterfly's inner twiddles
template <typename T, bool INV> __device__ __forceinline__ c2<T> w16(c2<T> a, int m) {
    const T c1 = T(0.92387953251128675612818318939678828682);
    const T s1 = T(0.38268343236508977172845998403039886676);
    const T h = T(0.70710678118654752440084436210484903928);
    T wr, wi;
    switch (m) {
        case 1: wr = c1; wi = -s1; break;
        case 2: wr = h; wi = -h; break;
        case 3: wr = s1; wi = -c1; break;
        case 4: wr = 0; wi = -1; break;
        case 6: wr = -h; wi = -h; break;
        case 9: wr = -c1; wi = s1; break;
        default: wr = 1; wi = 0; break;
    }
    if (INV) wi = -wi;
    return c2<T>{a.x * wr - a.y * wi, a.x * wi + a.y * wr};
}

template <typename T, bool INV> __device__ __forceinline__ void dft16(c2<T>* x) {
    c2<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        c2<T> t[4] = {x[n2], x[4 + n2], x[8 + n2], x[12 + n2]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = (n2 && k1) ? w16<T, INV>(t[k1], n2 * k1) : t[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        c2<T> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) x[k1 + 4 * k2] = t[k2];
    }
}

template <typename T, bool INV, int R> __device__ __forceinline__ void rdft(c2<T>* a) {
    if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 8) dft8<T, INV>(a);
    else dft16<T, INV>(a);
}

// butterflies of one stage on the registers (twiddles from the global table)
template <typename T, int E, int R, bool INV>
__device__ __forceinline__ void reg_bfly(c2<T> (&v)[E], const c2<T>* __restrict__ tw, int P, int Ns, int t, int Tl) {
    constexpr int Q = E / R;
    const int ts = P / (Ns * R);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (Ns > 1) {
            const int k = (t + Tl * q) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const c2<T> w = tw[r * k * ts];
                v[q * R + r] = INV ? cmulc(v[q * R + r], w) : cmul(v[q * R + r], w);
            }
        }
        rdft<T, INV, R>(&v[q * R]);
    }
}

template <typename T, int E, bool INV>
__device__ __forceinline__ void reg_bfly_any(c2<T> (&v)[E], const c2<T>* tw, int R, int P, int Ns, int t, int Tl) {
    if (R == 2) reg_bfly<T, E, 2, INV>(v, tw, P, Ns, t, Tl);
    else if (R == 4) { if constexpr (E % 4 == 0) reg_bfly<T, E, 4, INV>(v, tw, P, Ns, t, Tl); }
    else if (R == 8) { if constexpr (E % 8 == 0) reg_bfly<T, E, 8, INV>(v, tw, P, Ns, t, Tl); }
    else { if constexpr (E % 16 == 0) reg_bfly<T, E, 16, INV>(v, tw, P, Ns, t, Tl); }
}

// LDS exchange between stages: outputs of a stage (radix Rp, sub-size Nsp) ->
// inputs of the next stage (radix R). lds(pos) = lds[pos * sp + off].
template <typename T, int E>
__device__ __forceinline__ void reg_exchange(c2<T> (&v)[E], c2<T>* lds, int sp, int off, int P, int Rp, int Nsp,
                                             int R, int t, int Tl) {
    const int Qp = E / Rp;
    __syncthreads();   // previous readers of the buffer are done
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int q = m / Rp, r = m - q * Rp;
        if (q < Qp) {
            const int jj = t + Tl * q;
            const int k = jj & (Nsp - 1);
            const int o = (jj - k) * Rp + k + r * Nsp;
            lds[o * sp + off] = v[m];
        }
    }
    __syncthreads();
    const int Q = E / R;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int q = m / R, r = m - q * R;
        if (q < Q) v[m] = lds[((t + Tl * q) + r * (P / R)) * sp + off];
    }
}

// Full transform. On entry v holds the first stage's input layout, on exit
// the last stage's output layout (see header).
template <typename T, int E, bool INV>
__device__ __forceinline__ void reg_fft(c2<T> (&v)[E], c2<T>* lds, int sp, int off, const c2<T>* tw,
                                        const RegPlan& pl, int t) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.radix[s];
        if (s > 0) reg_exchange<T, E>(v, lds, sp, off, pl.P, pl.radix[s - 1], Ns / pl.radix[s - 1], R, t, pl.Tl);
        reg_bfly_any<T, E, INV>(v, tw, R, pl.P, Ns, t, pl.Tl);
        Ns *= R;
    }
}

// position of slot m for the first-stage input / last-stage output layout
__device__ __forceinline__ int reg_pos(int m, int R, int E, int t, int Tl) {
    const int q = m / R, r = m - q * R;
    return t + Tl * (q + (E / R) * r);
}

}  // namespace marie
