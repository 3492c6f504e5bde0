// The reference's default self-terms: DIRECTFN for constant basis functions
// and the weakly singular kernel exp(-j ko R) / R (Polimeridis & Mosig, fully
// numerical direct evaluation of 4-D singular integrals over triangle pairs;
// the reference's directfn_ws_st_const.cpp, directfn_ws_ea_const.cpp,
// directfn_ws_va_const.cpp). getOPERATORS.m:40-42 selects it by default, and
// singular_st.m:15-18, singular_ea.m:17-28 and singular_va.m:20-31 sum 20
// triangle-pair integrals into the five values I_ST, I_EAc, I_EAo, I_VAc,
// I_VAo that nearby_correct.m places into the near-offset blocks.
//
// Each triangle is parametrised over the equilateral simplex (eta in [-1, 1],
// xi in [0, sqrt3 (1 - |eta|)]; barycentric weights ((sqrt3 (1 - eta) - xi),
// (sqrt3 (1 + eta) - xi)) / (2 sqrt3) and xi / sqrt3). The 4-D integral is
// split into angular subdomains and mapped to polar-type coordinates in which
// the 1/R singularity is cancelled by the Jacobian; every remaining direction
// takes the same Np-point Gauss-Legendre rule:
//   ST (coincident triangle): 3 sub-triangles x 8 subdomains in (psi, eta),
//      then lambda and rho -- Np^4 points each;
//   EA (edge-adjacent pair): 6 subdomains in (theta, psi), two signs B,
//      one or two eta segments, then lambda;
//   VA (vertex-adjacent pair): theta_p, theta_q, psi split at atan(Lq/Lp),
//      then lambda.
// The reference's gateways hand the edge- and vertex-adjacent routines the
// second triangle as (r4, r2, r1) and (r4, r5, r1) (directfn_ws_ea_const_mex.cpp:50,
// directfn_ws_va_const_mex.cpp:51); the same vertex order is used here, since
// the values the reference assembles with are the ones to reproduce.
//
// On the GPU: one thread per outer quadrature node (ST: kk, subdomain, psi,
// eta; EA: subdomain, theta, psi, B; VA: theta_p, theta_q, psi), which runs
// the inner sums; the host adds every call's node contributions in node order
// (deterministic) and scales by the area Jacobians.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace marie {

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);   // vie_assembly.hip

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kS3 = 1.7320508075688772;       // sqrt(3), correctly rounded
constexpr int kMaxNp = 64;

enum DfnKind { DFN_ST = 0, DFN_EA = 1, DFN_VA = 2 };

struct DfnCall {
    int kind;
    int first;          // first node slot of this call in the output
    double P[9];        // vertices p1, p2, p3 (x, y, z each)
    double Q[9];        // vertices q1, q2, q3 (ST: unused)
};

__host__ __device__ inline int64_t dfn_nodes(int kind, int np) {
    return kind == DFN_ST ? 24LL * np * np : kind == DFN_EA ? 12LL * np * np : (int64_t)np * np * np;
}

// point of triangle T at simplex coordinates (eta, xi)
__device__ __forceinline__ void tri_point(const double* T, double eta, double xi, double r[3]) {
    const double b1 = (kS3 * (1.0 - eta) - xi) / (2.0 * kS3);
    const double b2 = (kS3 * (1.0 + eta) - xi) / (2.0 * kS3);
    const double b3 = xi / kS3;
    for (int c = 0; c < 3; ++c) r[c] = b1 * T[c] + b2 * T[3 + c] + b3 * T[6 + c];
}

// exp(-j ko R) / R between a point of P and a point of Q
__device__ __forceinline__ c2<double> green(const double* P, double ep, double xp, const double* Q, double eq,
                                            double xq, double ko) {
    double a[3], b[3];
    tri_point(P, ep, xp, a);
    tri_point(Q, eq, xq, b);
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    const double R = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double sn, cs;
    sincos(ko * R, &sn, &cs);
    return c2<double>{cs / R, -sn / R};
}

__device__ __forceinline__ void cacc(c2<double>& s, double w, c2<double> v) {
    s.x += w * v.x;
    s.y += w * v.y;
}

// map (lo, hi) of Gauss node i: value and half-length
__device__ __forceinline__ double node(const double* z, int i, double lo, double hi, double& J) {
    J = 0.5 * (hi - lo);
    return J * z[i] + 0.5 * (hi + lo);
}

// ---------------------------------------------------------------- coincident
// sub-triangle kk (0, 1, 2) of the equilateral simplex: rotation by 0, +-120 degrees
__device__ __forceinline__ void subtri(int kk, double e, double x, double& eo, double& xo) {
    if (kk == 0) {
        eo = e;
        xo = x;
    } else if (kk == 1) {
        eo = 0.5 - 0.5 * e - 0.5 * kS3 * x;
        xo = 0.5 * kS3 + 0.5 * kS3 * e - 0.5 * x;
    } else {
        eo = -0.5 - 0.5 * e + 0.5 * kS3 * x;
        xo = 0.5 * kS3 - 0.5 * kS3 * e - 0.5 * x;
    }
}

__device__ c2<double> st_node(const double* P, int64_t t, int np, const double* z, const double* w, double ko) {
    const int ie = (int)(t % np);
    const int ip = (int)((t / np) % np);
    const int m = (int)((t / ((int64_t)np * np)) % 8);       // subdomain 0..7
    const int kk = (int)(t / (8LL * np * np));
    // psi: [0, pi/3] (subdomains 0, 4, 5), [pi/3, 2pi/3] (1, 6), [2pi/3, pi] (2, 3, 7)
    const double plo = (m == 0 || m == 4 || m == 5) ? 0.0 : (m == 1 || m == 6) ? kPi / 3.0 : 2.0 * kPi / 3.0;
    const double phi = (m == 0 || m == 4 || m == 5) ? kPi / 3.0 : (m == 1 || m == 6) ? 2.0 * kPi / 3.0 : kPi;
    double Jp, Je;
    const double psi = node(z, ip, plo, phi, Jp);
    const double sp = sin(psi), cp = cos(psi);
    double elo, ehi;
    if (m == 0 || m == 1) {
        elo = 0.0; ehi = 1.0;
    } else if (m == 2 || m == 3) {
        const double b1 = tan(kPi - psi) / kS3, a1 = (1.0 - b1) / (1.0 + b1);
        elo = m == 2 ? a1 : 0.0;
        ehi = m == 2 ? 1.0 : a1;
    } else if (m == 4 || m == 5) {
        const double b2 = tan(psi) / kS3, a2 = -(1.0 - b2) / (1.0 + b2);
        elo = m == 4 ? a2 : -1.0;
        ehi = m == 4 ? 0.0 : a2;
    } else {
        elo = -1.0; ehi = 0.0;
    }
    const double eta = node(z, ie, elo, ehi, Je);
    // upper limit of lambda
    double lmax;
    if (m == 0 || m == 4) lmax = (1.0 - eta) / cp;
    else if (m == 1 || m == 2) lmax = kS3 * (1.0 - eta) / sp;
    else if (m == 3 || m == 7) lmax = -(1.0 + eta) / cp;
    else lmax = kS3 * (1.0 + eta) / sp;
    c2<double> In = {0.0, 0.0};
    for (int il = 0; il < np; ++il) {
        double Jl, Jr;
        const double lam = node(z, il, 0.0, lmax, Jl);
        const double xps = lam * sp;
        double ep, xp;
        subtri(kk, eta, xps, ep, xp);
        c2<double> Ir = {0.0, 0.0};
        for (int ir = 0; ir < np; ++ir) {
            const double rho = node(z, ir, 0.0, lam, Jr);
            double eq, xq;
            subtri(kk, rho * cp + eta, -rho * sp + xps, eq, xq);
            cacc(Ir, w[ir] * rho, green(P, ep, xp, P, eq, xq, ko));
        }
        cacc(In, w[il] * Jr, Ir);
    }
    const double f = Jp * w[ip] * Je * w[ie] * sp * (0.5 * lmax);
    return c2<double>{f * In.x, f * In.y};
}

// ------------------------------------------------------------ edge-adjacent
__device__ c2<double> ea_lambda(const double* P, const double* Q, double psi, double th, double eta, double lmax, int sgn,
                                int np, const double* z, const double* w, double ko) {
    const double sp = sin(psi), cp = cos(psi), ct = cos(th), stt = sin(th);
    c2<double> I = {0.0, 0.0};
    double Jl = 0.0;
    for (int il = 0; il < np; ++il) {
        const double lam = node(z, il, 0.0, lmax, Jl);
        const double ep = sgn * eta, xp = lam * sp;
        const double eq = sgn * (lam * cp * ct - eta), xq = lam * cp * stt;
        cacc(I, w[il] * lam * lam, green(P, ep, xp, Q, eq, xq, ko));
    }
    return c2<double>{Jl * I.x, Jl * I.y};
}

__device__ c2<double> ea_node(const double* P, const double* Q, int64_t t, int np, const double* z, const double* w,
                              double ko) {
    const int sgn = (t & 1) ? -1 : 1;          // B = 1, 2
    const int64_t u = t >> 1;
    const int ip = (int)(u % np);
    const int it = (int)((u / np) % np);
    const int m = (int)(u / ((int64_t)np * np));   // subdomain 0..5
    const double tlo = (m == 0 || m == 4) ? 0.0 : (m == 1 || m == 2) ? kPi / 3.0 : kPi / 2.0;
    const double thi = m == 0 ? kPi / 3.0 : (m == 1 || m == 2 || m == 4) ? kPi / 2.0 : kPi;
    double Jt, Jp;
    const double th = node(z, it, tlo, thi, Jt);
    const double st_ = sin(th), ct = cos(th);
    const double pa = atan(st_ - kS3 * ct), pb = atan(st_ + kS3 * ct);
    const double plo = m == 1 ? pa : m == 4 ? pb : m == 5 ? pa : 0.0;
    const double phi = (m == 0 || m == 1) ? pb : (m == 2 || m == 3) ? pa : kPi / 2.0;
    const double psi = node(z, ip, plo, phi, Jp);
    const double sp = sin(psi), cp = cos(psi);
    // lambda limits: L1 over (sin th - sqrt3 cos th), L2 over sin psi, L3 over (sin th + sqrt3 cos th)
    auto L1 = [&](double e) { return kS3 * (1.0 - fabs(e)) / (cp * (st_ - kS3 * ct)); };
    auto L2 = [&](double e) { return kS3 * (1.0 - fabs(e)) / sp; };
    auto L3 = [&](double e) { return kS3 * (1.0 + e) / (cp * (st_ + kS3 * ct)); };
    c2<double> N = {0.0, 0.0};
    // eta segments [0, s] with the first limit and [s, 1] with the second, or [0, 1] with one limit
    auto seg = [&](double lo, double hi, int lim) {
        double Je;
        c2<double> S = {0.0, 0.0};
        for (int ie = 0; ie < np; ++ie) {
            const double eta = node(z, ie, lo, hi, Je);
            const double lmax = lim == 1 ? L1(eta) : lim == 2 ? L2(eta) : L3(eta);
            cacc(S, w[ie], ea_lambda(P, Q, psi, th, eta, lmax, sgn, np, z, w, ko));
        }
        N.x += Je * S.x;
        N.y += Je * S.y;
    };
    if (m == 0 || m == 1) {
        const double g = tan(psi) / (st_ + kS3 * ct);
        const double s = (1.0 - g) / (1.0 + g);
        seg(0.0, s, 3);
        seg(s, 1.0, 2);
    } else if (m == 2) {
        const double s = kS3 / tan(th);
        seg(0.0, s, 3);
        seg(s, 1.0, 1);
    } else if (m == 3) {
        seg(0.0, 1.0, 1);
    } else {
        seg(0.0, 1.0, 2);
    }
    const double f = Jt * w[it] * Jp * w[ip] * cp;
    return c2<double>{f * N.x, f * N.y};
}

// ---------------------------------------------------------- vertex-adjacent
__device__ c2<double> va_node(const double* P, const double* Q, int64_t t, int np, const double* z, const double* w,
                              double ko) {
    const int ip = (int)(t % np);
    const int iq = (int)((t / np) % np);
    const int it = (int)(t / ((int64_t)np * np));
    double Jtp, Jtq, J1, J2, Jl;
    const double tp = node(z, it, 0.0, kPi / 3.0, Jtp);
    const double tq = node(z, iq, 0.0, kPi / 3.0, Jtq);
    const double Lp = 2.0 * kS3 / (sin(tp) + kS3 * cos(tp));
    const double Lq = 2.0 * kS3 / (sin(tq) + kS3 * cos(tq));
    const double pm = atan(Lq / Lp);
    const double ctp = cos(tp), stp = sin(tp), ctq = cos(tq), stq = sin(tq);
    auto part = [&](double psi, double lmax) {
        const double sp = sin(psi), cp = cos(psi);
        c2<double> I = {0.0, 0.0};
        for (int il = 0; il < np; ++il) {
            const double lam = node(z, il, 0.0, lmax, Jl);
            const double ep = lam * cp * ctp - 1.0, xp = lam * cp * stp;
            const double eq = lam * sp * ctq - 1.0, xq = lam * sp * stq;
            cacc(I, w[il] * (lam * lam * lam) * sp * cp, green(P, ep, xp, Q, eq, xq, ko));
        }
        return c2<double>{Jl * I.x, Jl * I.y};
    };
    const double psi1 = node(z, ip, 0.0, pm, J1);
    const c2<double> I1 = part(psi1, Lp / cos(psi1));
    const double psi2 = node(z, ip, pm, kPi / 2.0, J2);
    const c2<double> I2 = part(psi2, Lq / sin(psi2));
    const double f = Jtp * w[it] * Jtq * w[iq] * w[ip];
    return c2<double>{f * (J1 * I1.x + J2 * I2.x), f * (J1 * I1.y + J2 * I2.y)};
}

__global__ __launch_bounds__(256) void k_directfn(const DfnCall* __restrict__ calls, int np,
                                                  const double* __restrict__ zw, double ko,
                                                  c2<double>* __restrict__ out) {
    const DfnCall& c = calls[blockIdx.y];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dfn_nodes(c.kind, np)) return;
    const double* z = zw;
    const double* w = zw + np;
    c2<double> v;
    if (c.kind == DFN_ST) v = st_node(c.P, t, np, z, w, ko);
    else if (c.kind == DFN_EA) v = ea_node(c.P, c.Q, t, np, z, w, ko);
    else v = va_node(c.P, c.Q, t, np, z, w, ko);
    out[c.first + t] = v;
}

// area of triangle (a, b, c) / sqrt3: the Jacobian of the equilateral parametrisation
double tri_jac(const double* a, const double* b, const double* c) {
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], zz = u[0] * v[1] - u[1] * v[0];
    return 0.5 * std::sqrt(x * x + y * y + zz * zz) / std::sqrt(3.0);
}

}  // namespace

// The five self-terms as singular_st/ea/va.m compute them with method
// 'DIRECTFN' and Np_1D = np (assembly_nop.m:16 uses 20), into out[5].
int directfn_terms(double dx, double ko, int np, c2<double>* out) {
    const double d = dx;
    typedef std::array<double, 3> V;
    const V r1{0, 0, 0}, r2{d, 0, 0}, r3{d, d, 0}, r4{0, d, 0};
    std::vector<DfnCall> calls;
    std::vector<int> term;
    auto put = [](double* T, const V& a, const V& b, const V& c) {
        for (int k = 0; k < 3; ++k) { T[k] = a[k]; T[3 + k] = b[k]; T[6 + k] = c[k]; }
    };
    // the gateways' argument orders: st(a,b,c): P = (a,b,c); ea(a,b,c,e): P = (a,b,c),
    // Q = (e,b,a); va(a,b,c,e,f): P = (a,b,c), Q = (e,f,a)
    auto st = [&](int tm, const V& a, const V& b, const V& c) {
        DfnCall k{};
        k.kind = DFN_ST;
        put(k.P, a, b, c);
        calls.push_back(k);
        term.push_back(tm);
    };
    auto ea = [&](int tm, const V& a, const V& b, const V& c, const V& e) {
        DfnCall k{};
        k.kind = DFN_EA;
        put(k.P, a, b, c);
        put(k.Q, e, b, a);
        calls.push_back(k);
        term.push_back(tm);
    };
    auto va = [&](int tm, const V& a, const V& b, const V& c, const V& e, const V& f) {
        DfnCall k{};
        k.kind = DFN_VA;
        put(k.P, a, b, c);
        put(k.Q, e, f, a);
        calls.push_back(k);
        term.push_back(tm);
    };
    // I_ST (singular_st.m:15-18)
    st(0, r1, r2, r3);
    ea(0, r1, r3, r2, r4);
    ea(0, r1, r3, r4, r2);
    st(0, r1, r3, r4);
    // I_EAc, I_EAo (singular_ea.m:17-20, :25-28)
    const V r5{2 * d, 0, 0}, r6{2 * d, d, 0}, r5o{d, 0, d}, r6o{d, d, d};
    ea(1, r2, r3, r1, r5);
    va(1, r3, r1, r2, r5, r6);
    va(1, r3, r4, r1, r2, r5);
    va(1, r3, r4, r1, r5, r6);
    ea(2, r2, r3, r1, r5o);
    va(2, r3, r1, r2, r5o, r6o);
    va(2, r3, r4, r1, r2, r5o);
    va(2, r3, r4, r1, r5o, r6o);
    // I_VAc, I_VAo (singular_va.m:20-23, :28-31)
    const V s5{2 * d, d, 0}, s6{2 * d, 2 * d, 0}, s7{d, 2 * d, 0};
    const V s5o{d, d, d}, s6o{d, 2 * d, d}, s7o{d, 2 * d, 0};
    va(3, r3, r1, r2, s5, s6);
    va(3, r3, r1, r2, s6, s7);
    va(3, r3, r4, r1, s5, s6);
    va(3, r3, r4, r1, s6, s7);
    va(4, r3, r1, r2, s5o, s6o);
    va(4, r3, r1, r2, s6o, s7o);
    va(4, r3, r4, r1, s5o, s6o);
    va(4, r3, r4, r1, s6o, s7o);

    int64_t tot = 0, maxn = 0;
    for (auto& k : calls) {
        k.first = (int)tot;
        const int64_t n = dfn_nodes(k.kind, np);
        tot += n;
        maxn = std::max(maxn, n);
    }
    std::vector<double> x, w;
    gauss_legendre(np, x, w);
    std::vector<double> zw(2 * (size_t)np);
    for (int i = 0; i < np; ++i) { zw[i] = x[i]; zw[np + i] = w[i]; }
    DfnCall* dcalls = nullptr;
    double* dzw = nullptr;
    c2<double>* dout = nullptr;
    auto cleanup = [&]() {
        if (dcalls) (void)hipFree(dcalls);
        if (dzw) (void)hipFree(dzw);
        if (dout) (void)hipFree(dout);
    };
#define DTRY(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); set_error(#x " failed"); cleanup(); return ERR_HIP; } } while (0)
    DTRY(hipMalloc(&dcalls, calls.size() * sizeof(DfnCall)));
    DTRY(hipMalloc(&dzw, zw.size() * sizeof(double)));
    DTRY(hipMalloc(&dout, (size_t)tot * sizeof(c2<double>)));
    DTRY(hipMemcpy(dcalls, calls.data(), calls.size() * sizeof(DfnCall), hipMemcpyHostToDevice));
    DTRY(hipMemcpy(dzw, zw.data(), zw.size() * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_directfn, dim3((unsigned)((maxn + 255) / 256), (unsigned)calls.size()), dim3(256), 0, 0, dcalls,
                       np, dzw, ko, dout);
    DTRY(hipGetLastError());
    std::vector<c2<double>> h((size_t)tot);
    DTRY(hipMemcpy(h.data(), dout, (size_t)tot * sizeof(c2<double>), hipMemcpyDeviceToHost));
#undef DTRY
    cleanup();
    for (int t = 0; t < 5; ++t) out[t] = c2<double>{0.0, 0.0};
    for (size_t c = 0; c < calls.size(); ++c) {
        const DfnCall& k = calls[c];
        double sx = 0.0, sy = 0.0;
        const int64_t n = dfn_nodes(k.kind, np);
        for (int64_t i = 0; i < n; ++i) {
            sx += h[(size_t)(k.first + i)].x;
            sy += h[(size_t)(k.first + i)].y;
        }
        const double J = k.kind == DFN_ST ? tri_jac(k.P, k.P + 3, k.P + 6) * tri_jac(k.P, k.P + 3, k.P + 6)
                                          : tri_jac(k.P, k.P + 3, k.P + 6) * tri_jac(k.Q, k.Q + 3, k.Q + 6);
        out[term[c]].x += J * sx;
        out[term[c]].y += J * sy;
    }
    for (int t = 0; t < 5; ++t)
        if (!std::isfinite(out[t].x) || !std::isfinite(out[t].y)) {
            set_error("DIRECTFN self-terms are not finite");
            return ERR_NONFINITE;
        }
    return OK;
}

}  // namespace marie

extern "C" int marie_vie_singular_terms_directfn(double dx, double ko, int np, double* out10) {
    using namespace marie;
    if (!out10 || !(dx > 0) || !std::isfinite(ko) || np < 1 || np > kMaxNp) {
        set_error("invalid arguments (dx > 0, finite ko, 1 <= Np <= 64)");
        return MARIE_ERR_ARG;
    }
    c2<double> v[5];
    const int st = directfn_terms(dx, ko, np, v);
    if (st) return st;
    for (int i = 0; i < 5; ++i) { out10[2 * i] = v[i].x; out10[2 * i + 1] = v[i].y; }
    return MARIE_OK;
}
