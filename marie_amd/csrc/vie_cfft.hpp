// Register-resident Stockham FFT with a compile-time plan (P a power of two,
// or 2^a times 3, 5 or 7 with the odd factor dividing E).
//
// Same data movement as vie_regfft.hpp (thread t of Tl = P/E owns butterflies
// jj = t + Tl*q of each stage, slot v[q*R + r] holds input jj + r*P/R; LDS only
// between stages), but the line length P, the points per thread E and the
// radix sequence are template constants, so every slot index, twiddle index
// and LDS address is a shift/add of t and the butterflies are straight-line
// code on fixed registers.
#pragma once
#include "vie_regfft.hpp"

namespace marie {

// Raw buffer access to an array of at most 2^31 - 1 bytes: an offset at or
// beyond the range reads zero without touching memory, which replaces a
// branch around loads of padding (k >= N) or of columns past the grid.
constexpr int kBufOob = (int)0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// AUX: the instruction's cache-policy bits (0 default, 2 non-temporal)
template <typename T, int AUX = 0> __device__ __forceinline__ c2<T> buf_load(__amdgpu_buffer_rsrc_t rs, int off);
template <> __device__ __forceinline__ c2<double> buf_load<double, 0>(__amdgpu_buffer_rsrc_t rs, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    return c2<double>{__builtin_bit_cast(double, ((unsigned long long)(unsigned)v[1] << 32) | (unsigned)v[0]),
                      __builtin_bit_cast(double, ((unsigned long long)(unsigned)v[3] << 32) | (unsigned)v[2])};
}
template <> __device__ __forceinline__ c2<float> buf_load<float, 0>(__amdgpu_buffer_rsrc_t rs, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
    return c2<float>{__builtin_bit_cast(float, (unsigned)v[0]), __builtin_bit_cast(float, (unsigned)v[1])};
}
template <> __device__ __forceinline__ c2<double> buf_load<double, 2>(__amdgpu_buffer_rsrc_t rs, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 2);
    return c2<double>{__builtin_bit_cast(double, ((unsigned long long)(unsigned)v[1] << 32) | (unsigned)v[0]),
                      __builtin_bit_cast(double, ((unsigned long long)(unsigned)v[3] << 32) | (unsigned)v[2])};
}
template <> __device__ __forceinline__ c2<float> buf_load<float, 2>(__amdgpu_buffer_rsrc_t rs, int off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 2);
    return c2<float>{__builtin_bit_cast(float, (unsigned)v[0]), __builtin_bit_cast(float, (unsigned)v[1])};
}

// offset `lin` when `ok`, else beyond the range (no branch)
__device__ __forceinline__ int buf_off(int lin, bool ok) {
    const int m = -(int)ok;
    return (lin & m) | (kBufOob & ~m);
}

struct CPlanData {
    int n;
    int r[12];
};

// greedy radices from {16, 8, 4, 2} that divide E and what is left of P,
// then 3, 5, 7 for the odd factors (mixed-radix sides such as 192 = 4*4*4*3)
constexpr CPlanData make_cplan(int P, int E) {
    CPlanData d{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    int rem = P;
    while (rem > 1 && d.n < 12) {
        int R = 16;
        while (R > 1 && (R > E || E % R != 0 || rem % R != 0)) R >>= 1;
        if (R == 1) {
            R = (E % 3 == 0 && rem % 3 == 0) ? 3 : (E % 5 == 0 && rem % 5 == 0) ? 5
              : (E % 7 == 0 && rem % 7 == 0) ? 7 : 1;
            if (R == 1) break;
        }
        d.r[d.n] = R;
        d.n += 1;
        rem /= R;
    }
    return d;
}

// the plan's radices multiply to P (every instantiated (P, E) must have one)
constexpr bool cplan_ok(int P, int E) {
    int prod = 1;
    for (int i = 0; i < make_cplan(P, E).n; ++i) prod *= make_cplan(P, E).r[i];
    return prod == P && P % E == 0;
}

constexpr int cplan_radix(int P, int E, bool rev, int s) {
    return rev ? make_cplan(P, E).r[make_cplan(P, E).n - 1 - s] : make_cplan(P, E).r[s];
}

// product of the radices before stage s
constexpr int cplan_ns(int P, int E, bool rev, int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= cplan_radix(P, E, rev, i);
    return ns;
}

constexpr int cplan_first(int P, int E, bool rev) { return cplan_radix(P, E, rev, 0); }
constexpr int cplan_last(int P, int E, bool rev) { return cplan_radix(P, E, rev, make_cplan(P, E).n - 1); }

template <typename T, int P, int E, int R, int Ns, bool INV>
__device__ __forceinline__ void c_bfly(c2<T> (&v)[E], const c2<T>* __restrict__ tw, int t) {
    static_assert(cplan_ok(P, E), "no radix plan for this line length and points per thread");
    constexpr int Tl = P / E;
    constexpr int Q = E / R;
    constexpr int ts = P / (Ns * R);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if constexpr (Ns > 1) {
            const int k = (int)((unsigned)(t + Tl * q) % (unsigned)Ns);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const c2<T> w = tw[r * k * ts];
                v[q * R + r] = INV ? cmulc(v[q * R + r], w) : cmul(v[q * R + r], w);
            }
        }
        if constexpr (R == 3) dft3<T, INV>(&v[q * R]);
        else if constexpr (R == 5) dft5<T, INV>(&v[q * R]);
        else if constexpr (R == 7) dft_odd<T, INV, 7>(&v[q * R], tw, P);   // P/7-strided entries of the table
        else rdft<T, INV, R>(&v[q * R]);
    }
}

// LDS slot of exchange position o. PAD inserts one element after every 8:
// when consecutive lanes hold consecutive butterflies of one line (sp = 1),
// a stage of radix Rp writes at a stride of Rp elements, which without the
// padding lands every lane of a quarter-wave in the same banks.
template <bool PAD>
__device__ __forceinline__ int c_slot(int o) {
    return PAD ? o + (o >> 3) : o;
}

// elements of exchange area one line needs
template <int P, bool PAD>
constexpr int c_line_len() {
    return PAD ? P + P / 8 : P;
}

// outputs of a stage (radix Rp, sub-size Nsp) -> inputs of the next (radix R)
template <typename T, int P, int E, int Rp, int Nsp, int R, bool PAD>
__device__ __forceinline__ void c_exchange(c2<T> (&v)[E], c2<T>* lds, int sp, int off, int t) {
    constexpr int Tl = P / E;
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int q = m / Rp, r = m % Rp;
        const int jj = t + Tl * q;
        const int k = (int)((unsigned)jj % (unsigned)Nsp);
        const int o = (jj - k) * Rp + k + r * Nsp;
        lds[c_slot<PAD>(o) * sp + off] = v[m];
    }
    lds_barrier();
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int q = m / R, r = m % R;
        v[m] = lds[c_slot<PAD>((t + Tl * q) + r * (P / R)) * sp + off];
    }
}

template <typename T, int P, int E, bool INV, bool REV, int S, bool PAD>
__device__ __forceinline__ void c_fft_rec(c2<T> (&v)[E], c2<T>* lds, int sp, int off, const c2<T>* tw, int t) {
    constexpr int n = make_cplan(P, E).n;
    if constexpr (S < n) {
        constexpr int R = cplan_radix(P, E, REV, S);
        constexpr int Ns = cplan_ns(P, E, REV, S);
        if constexpr (S > 0) {
            constexpr int Rp = cplan_radix(P, E, REV, S - 1);
            c_exchange<T, P, E, Rp, Ns / Rp, R, PAD>(v, lds, sp, off, t);
        }
        c_bfly<T, P, E, R, Ns, INV>(v, tw, t);
        c_fft_rec<T, P, E, INV, REV, S + 1, PAD>(v, lds, sp, off, tw, t);
    }
}

// Full transform (REV: radices in reverse order, used so an inverse can start
// from the layout a forward transform ends in; PAD: padded exchange slots,
// lines c_line_len<P, true>() elements apart).
template <typename T, int P, int E, bool INV, bool REV, bool PAD = false>
__device__ __forceinline__ void c_fft(c2<T> (&v)[E], c2<T>* lds, int sp, int off, const c2<T>* tw, int t) {
    c_fft_rec<T, P, E, INV, REV, 0, PAD>(v, lds, sp, off, tw, t);
}

// position held by slot m in a first-stage input / last-stage output layout of radix R
template <int P, int E, int R>
__device__ __forceinline__ int c_pos(int m, int t) {
    constexpr int Tl = P / E;
    return t + Tl * ((m / R) + (E / R) * (m % R));
}

}  // namespace marie
