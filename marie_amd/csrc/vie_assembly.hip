// Device assembly of the N-operator circulant fN (getOPERATORS(r,f,'N')).
//
// Follows src_vie/src_operators/assembly_nop.m:1-142 (quadrature orders
// 4 / 10 / 20 and the n_medium = 5 region), mexcubatures/cubature_nop.cpp
// (face-pair cubature of exp(-jkR)/R with the 21 unique pairs and their
// translation-symmetric completion), nearby_correct.m (self-term tables),
// vector_projection.m (here in closed form:
//   P(k,l,(a,b)) = s_k s_l [d(p_k,p_l) d(a,b) - d(p_k,b) d(a,p_l)]
// for outward normals s_k e_{p_k}), circulant_nop.m / gperiodic_coeff_nop.m
// (reflection signs: a component changes sign under reflection of axis q iff
// exactly one of its two indices is q) and fft_operator.m.
//
// Kernels:
//   k_far        one thread per offset, 4x4 rule per face, projected to G
//   k_pairs      one workgroup per (offset, unique face pair) with the 10x10
//                or 20x20 rule, deterministic block reduction -> J
//   k_singular   the five weakly singular self-terms (own polar / pyramid
//                quadrature, 24 points per direction), one workgroup each
//   k_fixup      completes J for the medium/near offsets, applies the
//                nearby correction, projects and overwrites G
//   k_circulant  embeds G in the Px x Py x Pz circulant (in the fN buffer)
//   then a forward 3D FFT of the six components in place.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

#include <cmath>
#include <vector>

namespace marie {

constexpr int AT = 256;

// unique (obs, src) face pairs, cubature_nop.cpp:27-30
__constant__ int c_obs[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
__constant__ int c_src[21] = {0, 1, 2, 3, 4, 5, 0, 2, 3, 4, 5, 2, 3, 4, 5, 2, 4, 5, 4, 5, 4};

struct Rule {            // tensor rule on [0,1]^2 (gauss_2d.m:1-19), up to 20x20
    int np;              // points per face
    const double* w;     // np
    const double* u;
    const double* v;
};

__device__ __forceinline__ void face_geom(int f, double dx, double o[3], double a[3], double b[3]) {
    // cube_faces.m:1-39: origin vertex, edge a = v2 - v1, edge b = v4 - v1
    const int p = f >> 1, side = f & 1;
    for (int q = 0; q < 3; ++q) { o[q] = 0.0; a[q] = 0.0; b[q] = 0.0; }
    o[p] = side ? dx : 0.0;
    const int ea = (p == 0) ? 1 : 0;          // -x/+x: a = y ; -y/+y and -z/+z: a = x
    const int eb = (p == 2) ? 1 : 2;          // -z/+z: b = y ; else b = z
    a[ea] = dx;
    b[eb] = dx;
}

// fills J[6][6] (complex) from the 21 unique values (cubature_nop.cpp:96-114)
__device__ __forceinline__ void sym_fill(c2<double> J[6][6]) {
    J[1][1] = J[0][0];
    J[2][0] = J[1][3]; J[2][1] = J[0][3];
    J[3][0] = J[1][2]; J[3][1] = J[0][2]; J[3][3] = J[2][2];
    J[4][0] = J[1][5]; J[4][1] = J[0][5]; J[4][2] = J[3][5]; J[4][3] = J[2][5];
    J[5][0] = J[1][4]; J[5][1] = J[0][4]; J[5][2] = J[3][4]; J[5][3] = J[2][4]; J[5][5] = J[4][4];
}

__device__ __forceinline__ double proj(int k, int l, int c) {
    const int A[6] = {0, 0, 0, 1, 1, 2}, Bc[6] = {0, 1, 2, 1, 2, 2};
    const int a = A[c], b = Bc[c];
    const int pk = k >> 1, pl = l >> 1;
    const double s = ((k & 1) ? 1.0 : -1.0) * ((l & 1) ? 1.0 : -1.0);
    return s * ((pk == pl && a == b ? 1.0 : 0.0) - (pk == b && a == pl ? 1.0 : 0.0));
}

// G (column-major L x M x N x 6) at voxel offset vox from J
__device__ __forceinline__ void project_store(c2<double> J[6][6], c2<double>* G, int64_t vox, int64_t nv, double sc) {
    for (int c = 0; c < 6; ++c) {
        double gr = 0.0, gi = 0.0;
        for (int k = 0; k < 6; ++k)
            for (int l = 0; l < 6; ++l) {
                const double p = proj(k, l, c);
                if (p != 0.0) { gr += p * J[k][l].x; gi += p * J[k][l].y; }
            }
        G[vox + c * nv] = c2<double>{gr * sc, gi * sc};
    }
}

__device__ __forceinline__ c2<double> green(double ko, double R) {
    double s, c;
    sincos(ko * R, &s, &c);
    return c2<double>{c / R, -s / R};
}

// far field: one thread per offset, 4x4 rule (assembly_nop.m:37-59)
__global__ __launch_bounds__(AT) void k_far(c2<double>* __restrict__ G, int L, int M, int N, double dx, double ko,
                                            Rule rule) {
    const int64_t nv = (int64_t)L * M * N;
    const int64_t t = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (t >= nv) return;
    const int mx = (int)(t % L), my = (int)((t / L) % M), mz = (int)(t / ((int64_t)L * M));
    const double rm[3] = {mx * dx, my * dx, mz * dx};
    c2<double> J[6][6];
    for (int i = 0; i < 21; ++i) {
        double op[3], ap[3], bp[3], oq[3], aq[3], bq[3];
        face_geom(c_obs[i], dx, op, ap, bp);
        face_geom(c_src[i], dx, oq, aq, bq);
        double sr = 0.0, si = 0.0;
        for (int ip = 0; ip < rule.np; ++ip) {
            double rp[3];
            for (int q = 0; q < 3; ++q) rp[q] = op[q] + ap[q] * rule.u[ip] + bp[q] * rule.v[ip] + rm[q];
            for (int iq = 0; iq < rule.np; ++iq) {
                double d2 = 0.0;
                for (int q = 0; q < 3; ++q) {
                    const double rq = oq[q] + aq[q] * rule.u[iq] + bq[q] * rule.v[iq];
                    const double d = rq - rp[q];
                    d2 += d * d;
                }
                const double Wt = rule.w[ip] * rule.w[iq];
                const c2<double> g = green(ko, sqrt(d2));
                sr += Wt * g.x;
                si += Wt * g.y;
            }
        }
        J[c_obs[i]][c_src[i]] = c2<double>{sr, si};
    }
    sym_fill(J);
    const double dx4 = dx * dx * dx * dx;
    project_store(J, G, t, nv, dx4 / (4.0 * M_PI));
}

// one workgroup per (task, unique pair): J_unique[task*21 + pair]
__global__ __launch_bounds__(AT) void k_pairs(c2<double>* __restrict__ Ju, const int* __restrict__ offs, double dx,
                                              double ko, Rule rule) {
    const int task = blockIdx.x / 21, i = blockIdx.x % 21;
    const double rm[3] = {offs[3 * task] * dx, offs[3 * task + 1] * dx, offs[3 * task + 2] * dx};
    double op[3], ap[3], bp[3], oq[3], aq[3], bq[3];
    face_geom(c_obs[i], dx, op, ap, bp);
    face_geom(c_src[i], dx, oq, aq, bq);
    const int np = rule.np;
    double sr = 0.0, si = 0.0;
    for (int e = threadIdx.x; e < np * np; e += AT) {
        const int ip = e / np, iq = e % np;
        double d2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double rp = op[q] + ap[q] * rule.u[ip] + bp[q] * rule.v[ip] + rm[q];
            const double rq = oq[q] + aq[q] * rule.u[iq] + bq[q] * rule.v[iq];
            const double d = rq - rp;
            d2 += d * d;
        }
        // coincident quadrature points only occur on singular pairs, which the
        // nearby correction replaces; keep them finite here
        if (d2 > 0.0) {
            const double Wt = rule.w[ip] * rule.w[iq];
            const c2<double> g = green(ko, sqrt(d2));
            sr += Wt * g.x;
            si += Wt * g.y;
        }
    }
    __shared__ double sh[2 * AT];
    sh[threadIdx.x] = sr;
    sh[AT + threadIdx.x] = si;
    __syncthreads();
    for (int o = AT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sh[threadIdx.x] += sh[threadIdx.x + o];
            sh[AT + threadIdx.x] += sh[AT + threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Ju[blockIdx.x] = c2<double>{sh[0], sh[AT]};
}

// ------------------------------------------------------------ K operator
// Face pairs of the far/medium K cubature and their signs (cubature_kop.cpp:29-36)
__constant__ int c_kobs[12] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5};
__constant__ int c_ksrc[12] = {0, 1, 0, 1, 2, 3, 2, 3, 4, 5, 4, 5};
__constant__ double c_ksign[12] = {1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1};

// (1 - exp(-jkR)(1 + jkR)) / R^3, evaluated as cubature_kop.cpp:94 writes it
__device__ __forceinline__ c2<double> kgreen(double ko, double R) {
    double sn, cs;
    sincos(ko * R, &sn, &cs);
    const double kr = ko * R;
    // exp(-jkR) (1 + jkR) = (cos + kR sin) + j (kR cos - sin)
    const double re = 1.0 - (cs + kr * sn), im = -(kr * cs - sn);
    return c2<double>{re / R / R / R, im / R / R / R};
}

// far formula for every offset, 4x4 rule (assembly_kop.m:27-41): one thread per
// offset, IK_c = sum over 12 pairs of sign * sum W G(R) R_c, R = r_obs - r_src
__global__ __launch_bounds__(AT) void k_kfar(c2<double>* __restrict__ G, int L, int M, int N, double dx, double ko,
                                             double coef, Rule rule) {
    const int64_t nv = (int64_t)L * M * N;
    const int64_t t = blockIdx.x * (int64_t)AT + threadIdx.x;
    if (t >= nv) return;
    const int mx = (int)(t % L), my = (int)((t / L) % M), mz = (int)(t / ((int64_t)L * M));
    const double rm[3] = {mx * dx, my * dx, mz * dx};
    double ar[3] = {0.0, 0.0, 0.0}, ai[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 12; ++i) {
        double op[3], ap[3], bp[3], oq[3], aq[3], bq[3];
        face_geom(c_kobs[i], dx, op, ap, bp);
        face_geom(c_ksrc[i], dx, oq, aq, bq);
        double pr[3] = {0.0, 0.0, 0.0}, pi[3] = {0.0, 0.0, 0.0};
        for (int ip = 0; ip < rule.np; ++ip) {
            double rp[3];
            for (int q = 0; q < 3; ++q) rp[q] = op[q] + ap[q] * rule.u[ip] + bp[q] * rule.v[ip] + rm[q];
            for (int iq = 0; iq < rule.np; ++iq) {
                double d[3], d2 = 0.0;
                for (int q = 0; q < 3; ++q) {
                    d[q] = rp[q] - (oq[q] + aq[q] * rule.u[iq] + bq[q] * rule.v[iq]);
                    d2 += d[q] * d[q];
                }
                // coincident points occur only for offsets the near rule replaces
                if (d2 > 0.0) {
                    const double Wt = rule.w[ip] * rule.w[iq];
                    const c2<double> g = kgreen(ko, sqrt(d2));
                    for (int q = 0; q < 3; ++q) {
                        pr[q] += Wt * g.x * d[q];
                        pi[q] += Wt * g.y * d[q];
                    }
                }
            }
        }
        for (int q = 0; q < 3; ++q) {
            ar[q] += c_ksign[i] * pr[q];
            ai[q] += c_ksign[i] * pi[q];
        }
    }
    for (int q = 0; q < 3; ++q) G[t + q * nv] = c2<double>{ar[q] * coef, ai[q] * coef};
}

// one workgroup per (task, pair): sum W G(R) R_c (3 complex) of one face pair;
// near = 0 takes the 12 far pairs, near = 1 the 21 unique pairs
__global__ __launch_bounds__(AT) void k_kpairs(c2<double>* __restrict__ Jk, const int* __restrict__ offs, double dx,
                                               double ko, Rule rule, int near) {
    const int npair = near ? 21 : 12;
    const int task = blockIdx.x / npair, i = blockIdx.x % npair;
    const double rm[3] = {offs[3 * task] * dx, offs[3 * task + 1] * dx, offs[3 * task + 2] * dx};
    double op[3], ap[3], bp[3], oq[3], aq[3], bq[3];
    face_geom(near ? c_obs[i] : c_kobs[i], dx, op, ap, bp);
    face_geom(near ? c_src[i] : c_ksrc[i], dx, oq, aq, bq);
    const int np = rule.np;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = threadIdx.x; e < np * np; e += AT) {
        const int ip = e / np, iq = e % np;
        double d[3], d2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double rp = op[q] + ap[q] * rule.u[ip] + bp[q] * rule.v[ip] + rm[q];
            const double rq = oq[q] + aq[q] * rule.u[iq] + bq[q] * rule.v[iq];
            d[q] = rp - rq;
            d2 += d[q] * d[q];
        }
        // coincident points only on the coincident-panel pairs, which
        // nearby_correct_kop.m sets to zero
        if (d2 > 0.0) {
            const double Wt = rule.w[ip] * rule.w[iq];
            const c2<double> g = kgreen(ko, sqrt(d2));
            for (int q = 0; q < 3; ++q) {
                acc[2 * q] += Wt * g.x * d[q];
                acc[2 * q + 1] += Wt * g.y * d[q];
            }
        }
    }
    __shared__ double sh[6][AT];
    for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = AT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) Jk[(int64_t)blockIdx.x * 3 + q] = c2<double>{sh[2 * q][0], sh[2 * q + 1][0]};
}

// medium offsets (tasks 0..nmed-1, far formula) and the 8 near offsets
// (cubature_kop_near_sym.m: completion :67-81, nearby_correct_kop.m
// coincident-panel zeroing, rows J(r,:) = s_r I(r,:,c_r), :120-131); one
// thread per task, then scaled into G.
__global__ void k_kfix(c2<double>* __restrict__ G, const c2<double>* __restrict__ Jmed,
                       const c2<double>* __restrict__ Jnear, const int* __restrict__ offs, int nmed, int L, int M,
                       int N, double coef) {
    const int task = blockIdx.x * blockDim.x + threadIdx.x;
    if (task >= nmed + 8) return;
    const int mx = offs[3 * task], my = offs[3 * task + 1], mz = offs[3 * task + 2];
    const int64_t nv = (int64_t)L * M * N;
    const int64_t vox = mx + (int64_t)L * (my + (int64_t)M * mz);
    double ir[3] = {0.0, 0.0, 0.0}, ii[3] = {0.0, 0.0, 0.0};
    if (task < nmed) {
        if (mx < 2 && my < 2 && mz < 2) return;      // replaced by the near rule
        for (int i = 0; i < 12; ++i)
            for (int q = 0; q < 3; ++q) {
                const c2<double> v = Jmed[((int64_t)task * 12 + i) * 3 + q];
                ir[q] += c_ksign[i] * v.x;
                ii[q] += c_ksign[i] * v.y;
            }
    } else {
        const int t8 = task - nmed;
        c2<double> I[6][6][3];
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b)
                for (int q = 0; q < 3; ++q) I[a][b][q] = c2<double>{0.0, 0.0};
        for (int i = 0; i < 21; ++i)
            for (int q = 0; q < 3; ++q) I[c_obs[i]][c_src[i]][q] = Jnear[((int64_t)t8 * 21 + i) * 3 + q];
        // completion by translation (cubature_kop_near_sym.m:67-81, as cubature_nop.cpp:96-114)
        const int fill[15][4] = {{1, 1, 0, 0}, {2, 0, 1, 3}, {2, 1, 0, 3}, {3, 0, 1, 2}, {3, 1, 0, 2},
                                 {3, 3, 2, 2}, {4, 0, 1, 5}, {4, 1, 0, 5}, {4, 2, 3, 5}, {4, 3, 2, 5},
                                 {5, 0, 1, 4}, {5, 1, 0, 4}, {5, 2, 3, 4}, {5, 3, 2, 4}, {5, 5, 4, 4}};
        for (int f = 0; f < 15; ++f)
            for (int q = 0; q < 3; ++q) I[fill[f][0]][fill[f][1]][q] = I[fill[f][2]][fill[f][3]][q];
        // coincident-panel entries set to zero (nearby_correct_kop.m:5-77)
        if (mx == 0 && my == 0 && mz == 0) {
            for (int a = 0; a < 6; ++a)
                for (int q = 0; q < 3; ++q) I[a][a][q] = c2<double>{0.0, 0.0};
        } else if (mx == 1 && my == 0 && mz == 0) {
            for (int q = 0; q < 3; ++q) I[0][1][q] = c2<double>{0.0, 0.0};
        } else if (mx == 0 && my == 1 && mz == 0) {
            for (int q = 0; q < 3; ++q) I[2][3][q] = c2<double>{0.0, 0.0};
        } else if (mx == 0 && my == 0 && mz == 1) {
            for (int q = 0; q < 3; ++q) I[4][5][q] = c2<double>{0.0, 0.0};
        }
        // IK_c = sum_r s_r (I(r, 2c+1, c_r) - I(r, 2c, c_r)), s = -,+ per axis, c_r = r/2
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 6; ++r) {
                const double sg = (r & 1) ? 1.0 : -1.0;
                const int cr = r >> 1;
                ir[c] += sg * (I[r][2 * c + 1][cr].x - I[r][2 * c][cr].x);
                ii[c] += sg * (I[r][2 * c + 1][cr].y - I[r][2 * c][cr].y);
            }
    }
    for (int q = 0; q < 3; ++q) G[vox + q * nv] = c2<double>{ir[q] * coef, ii[q] * coef};
}

// ----------------------------------------------------------- self-terms
// Five integrals I = int_Sp int_Sq exp(-jkR)/R over squares of side a, by the
// convolution reduction to the difference vector (see header of this file):
//  ST : 4  int_[0,a]^2 g(|t|)(a-t1)(a-t2)
//  EAc: 2 [int_[0,a]^2 g t1 (a-t2) + int_[a,2a]x[0,a] g (2a-t1)(a-t2)]
//  VAc: int_[0,a]^2 g t1 t2 + 2 int_[a,2a]x[0,a] g (2a-t1) t2 + int_[a,2a]^2 g (2a-t1)(2a-t2)
//  EAo: 2 int_[0,a]^3 (a-s) g(|(u,v,s)|)
//  VAo: int_[0,a]^3 s g + int_[0,a]^2 x [a,2a] (2a-s) g
// Sub-squares / sub-cubes touching the origin use polar / pyramid coordinates.
struct GL24 {
    double x[24], w[24];   // rule on [0,1]
};

__device__ __forceinline__ c2<double> cexpm(double k, double r) {   // exp(-j k r)
    double s, c;
    sincos(k * r, &s, &c);
    return c2<double>{c, -s};
}

// polar part over [0,a]^2: sum over both triangles of F(r,c,s) * exp(-jkr), F without the 1/r
__device__ c2<double> polar_part(int kind, double k, double a, const GL24& q, int tid, int nt) {
    double sr = 0.0, si = 0.0;
    for (int e = tid; e < 2 * 24 * 24; e += nt) {
        const int tri = e / 576, it = (e / 24) % 24, ir = e % 24;
        const double lo = tri ? M_PI / 4 : 0.0;
        const double th = lo + (M_PI / 4) * q.x[it];
        const double wt = (M_PI / 4) * q.w[it];
        double s, c;
        sincos(th, &s, &c);
        const double R = tri ? a / s : a / c;
        const double r = R * q.x[ir];
        const double wr = R * q.w[ir];
        double f;
        if (kind == 0) f = (a - r * c) * (a - r * s);        // ST
        else if (kind == 1) f = (r * c) * (a - r * s);       // EAc
        else f = (r * c) * (r * s);                          // VAc
        const c2<double> ex = cexpm(k, r);
        sr += wt * wr * f * ex.x;
        si += wt * wr * f * ex.y;
    }
    return c2<double>{sr, si};
}

__device__ c2<double> box2_part(int kind, double k, double a, const GL24& q, int tid, int nt) {
    // kind 0: [a,2a]x[0,a] (2a-t1)(a-t2); 1: [a,2a]x[0,a] (2a-t1) t2; 2: [a,2a]^2 (2a-t1)(2a-t2)
    double sr = 0.0, si = 0.0;
    for (int e = tid; e < 24 * 24; e += nt) {
        const int i1 = e / 24, i2 = e % 24;
        const double t1 = a + a * q.x[i1];
        const double t2 = (kind == 2 ? a : 0.0) + a * q.x[i2];
        const double W = a * q.w[i1] * a * q.w[i2];
        double f;
        if (kind == 0) f = (2 * a - t1) * (a - t2);
        else if (kind == 1) f = (2 * a - t1) * t2;
        else f = (2 * a - t1) * (2 * a - t2);
        const double r = sqrt(t1 * t1 + t2 * t2);
        const c2<double> ex = cexpm(k, r);
        sr += W * f * ex.x / r;
        si += W * f * ex.y / r;
    }
    return c2<double>{sr, si};
}

// pyramids over [0,a]^3 with apex at 0: g * Jacobian = a^2 lam exp(-jk a lam D)/D
__device__ c2<double> pyr_part(int kind, double k, double a, const GL24& q, int tid, int nt) {
    double sr = 0.0, si = 0.0;
    for (int e = tid; e < 24 * 24 * 24; e += nt) {
        const int il = e / 576, ip = (e / 24) % 24, iq = e % 24;
        const double lam = q.x[il], p = q.x[ip], qq = q.x[iq];
        const double W = q.w[il] * q.w[ip] * q.w[iq];
        const double D = sqrt(1.0 + p * p + qq * qq);
        const c2<double> ex = cexpm(k, a * lam * D);
        const double base = a * a * lam / D;
        double f;
        if (kind == 0) f = 2.0 * (2.0 * a * (1.0 - lam * qq) + a * (1.0 - lam));   // EAo, all pyramids
        else f = 2.0 * a * lam * qq + a * lam;                                       // VAo part A
        sr += W * base * f * ex.x;
        si += W * base * f * ex.y;
    }
    return c2<double>{sr, si};
}

__device__ c2<double> box3_part(double k, double a, const GL24& q, int tid, int nt) {
    // VAo part B: u,v in [0,a], s in [a,2a], weight (2a - s)
    double sr = 0.0, si = 0.0;
    for (int e = tid; e < 24 * 24 * 24; e += nt) {
        const int i1 = e / 576, i2 = (e / 24) % 24, i3 = e % 24;
        const double u = a * q.x[i1], v = a * q.x[i2], s = a + a * q.x[i3];
        const double W = a * q.w[i1] * a * q.w[i2] * a * q.w[i3];
        const double r = sqrt(u * u + v * v + s * s);
        const c2<double> ex = cexpm(k, r);
        sr += W * (2 * a - s) * ex.x / r;
        si += W * (2 * a - s) * ex.y / r;
    }
    return c2<double>{sr, si};
}

__global__ __launch_bounds__(AT) void k_singular(c2<double>* __restrict__ out, double ko, double a, GL24 q) {
    const int which = blockIdx.x;   // 0 ST, 1 EAc, 2 EAo, 3 VAc, 4 VAo
    c2<double> v = {0.0, 0.0};
    const int t = threadIdx.x;
    if (which == 0) {
        v = polar_part(0, ko, a, q, t, AT);
        v = c2<double>{4.0 * v.x, 4.0 * v.y};
    } else if (which == 1) {
        const c2<double> A = polar_part(1, ko, a, q, t, AT), B = box2_part(0, ko, a, q, t, AT);
        v = c2<double>{2.0 * (A.x + B.x), 2.0 * (A.y + B.y)};
    } else if (which == 2) {
        v = pyr_part(0, ko, a, q, t, AT);
    } else if (which == 3) {
        const c2<double> A = polar_part(2, ko, a, q, t, AT), B = box2_part(1, ko, a, q, t, AT),
                         C = box2_part(2, ko, a, q, t, AT);
        v = c2<double>{A.x + 2.0 * B.x + C.x, A.y + 2.0 * B.y + C.y};
    } else {
        const c2<double> A = pyr_part(1, ko, a, q, t, AT), B = box3_part(ko, a, q, t, AT);
        v = c2<double>{A.x + B.x, A.y + B.y};
    }
    __shared__ double sh[2 * AT];
    sh[t] = v.x;
    sh[AT + t] = v.y;
    __syncthreads();
    for (int o = AT / 2; o > 0; o >>= 1) {
        if (t < o) { sh[t] += sh[t + o]; sh[AT + t] += sh[AT + t + o]; }
        __syncthreads();
    }
    if (t == 0) out[which] = c2<double>{sh[0], sh[AT]};
}

// nearby_correct.m:5-133 tables, encoded per offset (0-based mx,my,mz -> 4mx+2my+mz... ):
// code: 0 keep computed value, 1 ST, 2 EAc, 3 EAo, 4 VAc, 5 VAo, 6 zero
__constant__ unsigned char c_near[8][6][6] = {
    // (0,0,0)
    {{1, 0, 3, 3, 3, 3}, {0, 1, 3, 3, 3, 3}, {3, 3, 1, 0, 3, 3}, {3, 3, 0, 1, 3, 3}, {3, 3, 3, 3, 1, 0}, {3, 3, 3, 3, 0, 1}},
    // (0,0,1)
    {{2, 0, 5, 5, 0, 3}, {0, 2, 5, 5, 0, 3}, {5, 5, 2, 0, 0, 3}, {5, 5, 0, 2, 0, 3}, {3, 3, 3, 3, 0, 1}, {0, 0, 0, 0, 0, 0}},
    // (0,1,0)
    {{2, 0, 0, 3, 5, 5}, {0, 2, 0, 3, 5, 5}, {3, 3, 0, 1, 3, 3}, {0, 0, 0, 0, 0, 0}, {5, 5, 0, 3, 2, 0}, {5, 5, 0, 3, 0, 2}},
    // (0,1,1)
    {{4, 0, 0, 5, 0, 5}, {0, 4, 0, 5, 0, 5}, {5, 5, 0, 2, 0, 3}, {0, 0, 0, 0, 0, 0}, {5, 5, 0, 3, 0, 2}, {0, 0, 0, 0, 0, 0}},
    // (1,0,0)
    {{0, 1, 3, 3, 3, 3}, {0, 0, 0, 0, 0, 0}, {0, 3, 2, 0, 5, 5}, {0, 3, 0, 2, 5, 5}, {0, 3, 5, 5, 2, 0}, {0, 3, 5, 5, 0, 2}},
    // (1,0,1)
    {{0, 2, 5, 5, 0, 3}, {0, 0, 0, 0, 0, 0}, {0, 5, 4, 0, 0, 5}, {0, 5, 0, 4, 0, 5}, {0, 3, 5, 5, 0, 2}, {0, 0, 0, 0, 0, 0}},
    // (1,1,0)
    {{0, 2, 0, 3, 5, 5}, {0, 0, 0, 0, 0, 0}, {0, 3, 0, 2, 5, 5}, {0, 0, 0, 0, 0, 0}, {0, 5, 0, 5, 4, 0}, {0, 5, 0, 5, 0, 4}},
    // (1,1,1)
    {{0, 4, 0, 5, 0, 5}, {0, 0, 0, 0, 0, 0}, {0, 5, 0, 4, 0, 5}, {0, 0, 0, 0, 0, 0}, {0, 5, 0, 5, 0, 4}, {0, 0, 0, 0, 0, 0}},
};
// P mask (1 = keep computed, 0 = use table value incl. zero), nearby_correct.m
__constant__ unsigned char c_keep[8][6][6] = {
    {{0, 1, 0, 0, 0, 0}, {1, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 0}, {0, 0, 1, 0, 0, 0}, {0, 0, 0, 0, 0, 1}, {0, 0, 0, 0, 1, 0}},
    {{0, 1, 0, 0, 1, 0}, {1, 0, 0, 0, 1, 0}, {0, 0, 0, 1, 1, 0}, {0, 0, 1, 0, 1, 0}, {0, 0, 0, 0, 1, 0}, {1, 1, 1, 1, 1, 1}},
    {{0, 1, 1, 0, 0, 0}, {1, 0, 1, 0, 0, 0}, {0, 0, 1, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {0, 0, 1, 0, 0, 1}, {0, 0, 1, 0, 1, 0}},
    {{0, 1, 1, 0, 1, 0}, {1, 0, 1, 0, 1, 0}, {0, 0, 1, 0, 1, 0}, {1, 1, 1, 1, 1, 1}, {0, 0, 1, 0, 1, 0}, {1, 1, 1, 1, 1, 1}},
    {{1, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 0, 1, 0, 0}, {1, 0, 1, 0, 0, 0}, {1, 0, 0, 0, 0, 1}, {1, 0, 0, 0, 1, 0}},
    {{1, 0, 0, 0, 1, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 0, 1, 1, 0}, {1, 0, 1, 0, 1, 0}, {1, 0, 0, 0, 1, 0}, {1, 1, 1, 1, 1, 1}},
    {{1, 0, 1, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 1, 0, 0, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 1, 0, 0, 1}, {1, 0, 1, 0, 1, 0}},
    {{1, 0, 1, 0, 1, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 1, 0, 1, 0}, {1, 1, 1, 1, 1, 1}, {1, 0, 1, 0, 1, 0}, {1, 1, 1, 1, 1, 1}},
};

// tasks: [0, nmed) medium offsets, [nmed, nmed+8) near offsets (4mx+2my+mz order)
__global__ void k_fixup(c2<double>* __restrict__ G, const c2<double>* __restrict__ Ju, const int* __restrict__ offs,
                        int nmed, const c2<double>* __restrict__ sing, int L, int M, int N, double dx) {
    const int task = blockIdx.x * blockDim.x + threadIdx.x;
    if (task >= nmed + 8) return;
    const int mx = offs[3 * task], my = offs[3 * task + 1], mz = offs[3 * task + 2];
    if (mx >= L || my >= M || mz >= N) return;
    // the near tasks own the [0,2)^3 offsets (assembly_nop.m:100-114 overwrite the medium values)
    if (task < nmed && mx < 2 && my < 2 && mz < 2) return;
    c2<double> J[6][6];
    for (int i = 0; i < 21; ++i) J[c_obs[i]][c_src[i]] = Ju[task * 21 + i];
    sym_fill(J);
    const double dx4 = dx * dx * dx * dx;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) J[i][j] = c2<double>{J[i][j].x * dx4, J[i][j].y * dx4};
    if (task >= nmed) {
        const int code = (mx << 2) | (my << 1) | mz;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) {
                if (!c_keep[code][i][j]) {
                    const int t = c_near[code][i][j];
                    J[i][j] = (t >= 1 && t <= 5) ? sing[t - 1] : c2<double>{0.0, 0.0};
                }
            }
    }
    const int64_t nv = (int64_t)L * M * N;
    project_store(J, G, mx + (int64_t)L * (my + (int64_t)M * mz), nv, 1.0 / (4.0 * M_PI));
}

// circulant embedding (circulant_nop.m:23-40) into fN buffer [6][Pz][Py][Px]
// kind 0: the 6 tensor components of N (signs of gperiodic_coeff_nop.m: a
// component flips once per reflected axis among its two indices); kind 1: the
// 3 components of K (gperiodic_coeff_kop.m: component c flips when axis c is
// reflected)
__global__ void k_circulant(const c2<double>* __restrict__ G, c2<double>* __restrict__ C, int L, int M, int N,
                            int Px, int Py, int Pz, int kind) {
    const int nc = kind ? 3 : 6;
    const int64_t tot = (int64_t)Px * Py * Pz * nc;
    const int A[6] = {0, 0, 0, 1, 1, 2}, Bc[6] = {0, 1, 2, 1, 2, 2};
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % Px);
        const int j = (int)((e / Px) % Py);
        const int k = (int)((e / ((int64_t)Px * Py)) % Pz);
        const int c = (int)(e / ((int64_t)Px * Py * Pz));
        int si, sj, sk;
        bool ri = false, rj = false, rk = false, zero = false;
        if (i < L) si = i; else if (i > Px - L) { si = Px - i; ri = true; } else { si = 0; zero = true; }
        if (j < M) sj = j; else if (j > Py - M) { sj = Py - j; rj = true; } else { sj = 0; zero = true; }
        if (k < N) sk = k; else if (k > Pz - N) { sk = Pz - k; rk = true; } else { sk = 0; zero = true; }
        c2<double> v = {0.0, 0.0};
        if (!zero) {
            v = G[si + (int64_t)L * (sj + (int64_t)M * (sk + (int64_t)N * c))];
            int flips = 0;
            if (kind) {
                flips = (ri && c == 0) + (rj && c == 1) + (rk && c == 2);
            } else {
                if (ri) flips += (A[c] == 0) + (Bc[c] == 0);
                if (rj) flips += (A[c] == 1) + (Bc[c] == 1);
                if (rk) flips += (A[c] == 2) + (Bc[c] == 2);
            }
            if (flips & 1) v = c2<double>{-v.x, -v.y};
        }
        C[e] = v;
    }
}

// contiguous-line forward FFT (x axis) of a full array, in place
template <int EMAX, int RSET>
__global__ __launch_bounds__(AT) void k_lines(c2<double>* __restrict__ data, const c2<double>* __restrict__ tw,
                                              AxisPlan pl, int64_t nlines, int lg) {
    extern __shared__ unsigned char smem_[];
    c2<double>* buf = reinterpret_cast<c2<double>*>(smem_);
    const int TX = 1 << lg;
    const int P = pl.P;
    const int64_t l0 = (int64_t)blockIdx.x * TX;
    for (int e = threadIdx.x; e < TX * P; e += AT) {
        const int b = e / P, i = e - b * P;
        c2<double> v = {0.0, 0.0};
        if (l0 + b < nlines) v = data[(l0 + b) * P + i];
        buf[i * TX + b] = v;
    }
    __syncthreads();
    tile_fft<double, EMAX, false, RSET>(buf, tw, pl, lg);
    for (int e = threadIdx.x; e < TX * P; e += AT) {
        const int b = e / P, i = e - b * P;
        if (l0 + b < nlines) data[(l0 + b) * P + i] = buf[i * TX + b];
    }
}

// strided-axis forward FFT of a full array, in place (tile of TX x positions)
template <int EMAX, int RSET>
__global__ __launch_bounds__(AT) void k_axis(c2<double>* __restrict__ data, const c2<double>* __restrict__ tw,
                                             AxisPlan pl, int Px, int64_t ls, int64_t plane_stride, int lg) {
    extern __shared__ unsigned char smem_[];
    c2<double>* buf = reinterpret_cast<c2<double>*>(smem_);
    const int TX = 1 << lg, msk = TX - 1;
    const int P = pl.P;
    const int i0 = blockIdx.x * TX;
    c2<double>* d = data + (int64_t)blockIdx.y * plane_stride + i0;
    const int nx = min(TX, Px - i0);
    for (int e = threadIdx.x; e < TX * P; e += AT) {
        const int b = e & msk, p = e >> lg;
        buf[e] = (b < nx) ? d[p * ls + b] : c2<double>{0.0, 0.0};
    }
    __syncthreads();
    tile_fft<double, EMAX, false, RSET>(buf, tw, pl, lg);
    for (int e = threadIdx.x; e < TX * P; e += AT) {
        const int b = e & msk, p = e >> lg;
        if (b < nx) d[p * ls + b] = buf[e];
    }
}

// ------------------------------------------------------------------ host side
// Gauss-Legendre on [-1,1], ascending, Newton in extended precision
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        long double z = cosl(3.141592653589793238462643383279502884L * (i + 0.75L) / (n + 0.5L));
        long double pp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                const long double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            const long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) {
                // one more evaluation for the derivative at the converged node
                p1 = 1; p2 = 0;
                for (int j = 1; j <= n; ++j) {
                    const long double p3 = p2;
                    p2 = p1;
                    p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
                }
                pp = n * (z * p1 - p2) / (z * z - 1);
                break;
            }
        }
        x[n - 1 - i] = (double)z;   // cos is decreasing in i: store ascending
        w[n - 1 - i] = (double)(2.0L / ((1 - z * z) * pp * pp));
    }
}

struct DevRule {
    double* buf = nullptr;
    Rule r{};
};

static int make_rule(int order, DevRule& dr) {
    std::vector<double> x, w;
    gauss_legendre(order, x, w);
    const int np = order * order;
    std::vector<double> h(3 * (size_t)np);
    int ctr = 0;
    for (int I = 0; I < order; ++I)
        for (int J = 0; J < order; ++J, ++ctr) {
            h[ctr] = w[I] * w[J] / 4.0;
            h[np + ctr] = (x[I] + 1.0) / 2.0;
            h[2 * np + ctr] = (x[J] + 1.0) / 2.0;
        }
    MARIE_HIP(hipMalloc(&dr.buf, h.size() * sizeof(double)));
    MARIE_HIP(hipMemcpy(dr.buf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    dr.r.np = np;
    dr.r.w = dr.buf;
    dr.r.u = dr.buf + np;
    dr.r.v = dr.buf + 2 * np;
    return OK;
}

static GL24 gl24() {
    std::vector<double> x, w;
    gauss_legendre(24, x, w);
    GL24 q;
    for (int i = 0; i < 24; ++i) {
        q.x[i] = (x[i] + 1.0) / 2.0;
        q.w[i] = w[i] / 2.0;
    }
    return q;
}

int singular_terms(double dx, double ko, c2<double>* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_singular, dim3(5), dim3(AT), 0, s, d_out, ko, dx, gl24());
    MARIE_HIP(hipGetLastError());
    return OK;
}

int fft3_forward_inplace(Op* h, c2<double>* data, int ncomp, hipStream_t s) {
    const c2<double>* tw = reinterpret_cast<const c2<double>*>(h->tw64);
    const int Px = h->Px, Py = h->Py, Pz = h->Pz;
    constexpr int EM = 16;
    const int cap = EM * AT;
    if (Px > cap || Py > cap || Pz > cap) { set_error("circulant side exceeds the single-tile FFT limit"); return ERR_SIZE; }
    auto lgfor = [&](int P, int txmax) { int lg = 0; while ((2 << lg) * P <= cap && (2 << lg) <= txmax) ++lg; return lg; };
    {
        const int lg = lgfor(Px, 16);
        const int64_t nl = (int64_t)Py * Pz * ncomp;
        MARIE_RSET(h->plan[0], hipLaunchKernelGGL((k_lines<EM, RS>), dim3((unsigned)((nl + (1 << lg) - 1) >> lg)), dim3(AT),
                           ((size_t)Px << lg) * sizeof(c2<double>), s, data, tw, h->plan[0], nl, lg));
        MARIE_HIP(hipGetLastError());
    }
    {
        const int lg = lgfor(Py, 16);
        MARIE_RSET(h->plan[1], hipLaunchKernelGGL((k_axis<EM, RS>), dim3((Px + (1 << lg) - 1) >> lg, Pz * ncomp), dim3(AT),
                           ((size_t)Py << lg) * sizeof(c2<double>), s, data, tw + Px, h->plan[1], Px, (int64_t)Px,
                           (int64_t)Px * Py, lg));
        MARIE_HIP(hipGetLastError());
    }
    {
        const int lg = lgfor(Pz, 16);
        for (int c = 0; c < ncomp; ++c) {
            MARIE_RSET(h->plan[2], hipLaunchKernelGGL((k_axis<EM, RS>), dim3((Px + (1 << lg) - 1) >> lg, Py), dim3(AT),
                               ((size_t)Pz << lg) * sizeof(c2<double>), s, data + (int64_t)c * Px * Py * Pz,
                               tw + Px + Py, h->plan[2], Px, (int64_t)Px * Py, (int64_t)Px, lg));
            MARIE_HIP(hipGetLastError());
        }
    }
    return OK;
}

__global__ void k_to_f32(const c2<double>* __restrict__ s, c2<float>* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = c2<float>{(float)s[i].x, (float)s[i].y};
}

// copy n complex of a device array to the host (stream-ordered, waited for)
static int to_host(const c2<double>* d, double* host, size_t n, hipStream_t s) {
    MARIE_HIP(hipMemcpyAsync(host, d, n * sizeof(c2<double>), hipMemcpyDeviceToHost, s));
    MARIE_HIP(hipStreamSynchronize(s));
    return OK;
}

// stage < 0: assemble and store fN on the handle (getOPERATORS 'FFTofC');
// stage 0 / 1: copy the Toeplitz generator (L x M x N x 6, 'T') or the
// circulant (Px x Py x Pz x 6, 'C') to `out` instead, leaving the handle's
// operator as it was (getOPERATORS.m:88-95)
int assemble_N(Op* h, double dx, double ko, const double* singular, int stage, double* out) {
    const int L = h->L, M = h->M, N = h->N;
    // any circulant side >= 2L-1 embeds the same Toeplitz operator (circulant_nop.m
    // uses exactly 2L); other sides let the caller pick FFT-friendly factors
    if (h->Px < 2 * L - 1 || h->Py < 2 * M - 1 || h->Pz < 2 * N - 1) {
        set_error("assembly needs a circulant of at least (2L-1) x (2M-1) x (2N-1)");
        return ERR_ARG;
    }
    if (L < 5 || M < 5 || N < 5) {
        set_error("assembly needs every grid side >= 5 (n_medium of assembly_nop.m)");
        return ERR_ARG;
    }
    if (!(dx > 0) || !std::isfinite(ko) || ko < 0) { set_error("invalid dx or ko"); return ERR_ARG; }
    hipStream_t s = h->stream;
    const int64_t nv = (int64_t)L * M * N;
    const size_t nf = (size_t)6 * h->Px * h->Py * h->Pz;
    void* full = nullptr;
    if (hipMalloc(&full, nf * sizeof(c2<double>)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for fN");
        return ERR_ALLOC;
    }
    c2<double>* G = nullptr;
    c2<double>* Ju = nullptr;
    c2<double>* sing = nullptr;
    int* offs = nullptr;
    DevRule r4, r10, r20;
    int st = OK;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        if (G) (void)hipFree(G);
        if (Ju) (void)hipFree(Ju);
        if (sing) (void)hipFree(sing);
        if (offs) (void)hipFree(offs);
        if (full) (void)hipFree(full);
        if (r4.buf) (void)hipFree(r4.buf);
        if (r10.buf) (void)hipFree(r10.buf);
        if (r20.buf) (void)hipFree(r20.buf);
    };
#define TRY(x) do { st = (x); if (st) { cleanup(); return st; } } while (0)
#define TRYH(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); set_error(#x " failed"); cleanup(); return ERR_HIP; } } while (0)
    TRYH(hipMalloc(&G, nv * 6 * sizeof(c2<double>)));
    const int nmed = 125;
    TRYH(hipMalloc(&Ju, (size_t)(nmed + 8) * 21 * sizeof(c2<double>)));
    TRYH(hipMalloc(&sing, 5 * sizeof(c2<double>)));
    TRYH(hipMalloc(&offs, (size_t)(nmed + 8) * 3 * sizeof(int)));
    std::vector<int> ho;
    for (int k = 0; k < 5; ++k)
        for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i) { ho.push_back(i); ho.push_back(j); ho.push_back(k); }
    for (int code = 0; code < 8; ++code) { ho.push_back(code >> 2); ho.push_back((code >> 1) & 1); ho.push_back(code & 1); }
    TRYH(hipMemcpyAsync(offs, ho.data(), ho.size() * sizeof(int), hipMemcpyHostToDevice, s));
    TRY(make_rule(4, r4));
    TRY(make_rule(10, r10));
    TRY(make_rule(20, r20));
    if (singular) {
        std::vector<c2<double>> sv(5);
        for (int i = 0; i < 5; ++i) sv[i] = c2<double>{singular[2 * i], singular[2 * i + 1]};
        TRYH(hipMemcpyAsync(sing, sv.data(), 5 * sizeof(c2<double>), hipMemcpyHostToDevice, s));
        TRYH(hipStreamSynchronize(s));
    } else {
        TRY(singular_terms(dx, ko, sing, s));
    }
    hipLaunchKernelGGL(k_far, dim3((unsigned)((nv + AT - 1) / AT)), dim3(AT), 0, s, G, L, M, N, dx, ko, r4.r);
    hipLaunchKernelGGL(k_pairs, dim3(nmed * 21), dim3(AT), 0, s, Ju, offs, dx, ko, r10.r);
    hipLaunchKernelGGL(k_pairs, dim3(8 * 21), dim3(AT), 0, s, Ju + nmed * 21, offs + 3 * nmed, dx, ko, r20.r);
    hipLaunchKernelGGL(k_fixup, dim3(1), dim3(192), 0, s, G, Ju, offs, nmed, sing, L, M, N, dx);
    if (stage == 0) {
        TRYH(hipGetLastError());
        TRY(to_host(G, out, (size_t)nv * 6, s));
        cleanup();
        return OK;
    }
    hipLaunchKernelGGL(k_circulant, dim3(4096), dim3(256), 0, s, G, (c2<double>*)full, L, M, N, h->Px, h->Py, h->Pz,
                       0);
    TRYH(hipGetLastError());
    if (stage == 1) {
        TRY(to_host((const c2<double>*)full, out, nf, s));
        cleanup();
        return OK;
    }
    TRY(fft3_forward_inplace(h, (c2<double>*)full, 6, s));
    TRYH(hipStreamSynchronize(s));
    {
        void* f = full;
        full = nullptr;          // ownership passes to the handle
        TRY(store_fN(h, f));
    }
#undef TRY
#undef TRYH
    cleanup();
    return OK;
}

// stage as assemble_N (getOPERATORS.m:101-115): 'T' is L x M x N x 3, 'C' Px x Py x Pz x 3
int assemble_K(Op* h, double dx, double ko, int stage, double* out) {
    const int L = h->L, M = h->M, N = h->N;
    if (h->Px < 2 * L - 1 || h->Py < 2 * M - 1 || h->Pz < 2 * N - 1) {
        set_error("assembly needs a circulant of at least (2L-1) x (2M-1) x (2N-1)");
        return ERR_ARG;
    }
    if (L < 5 || M < 5 || N < 5) {
        set_error("assembly needs every grid side >= 5 (n_medium of assembly_kop.m)");
        return ERR_ARG;
    }
    if (!(dx > 0) || !std::isfinite(ko) || !(ko > 0)) { set_error("invalid dx or ko (ko > 0)"); return ERR_ARG; }
    hipStream_t s = h->stream;
    const int64_t nv = (int64_t)L * M * N;
    const size_t nf = (size_t)3 * h->Px * h->Py * h->Pz;
    void* full = nullptr;
    if (hipMalloc(&full, nf * sizeof(c2<double>)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("out of device memory for fK");
        return ERR_ALLOC;
    }
    c2<double>* G = nullptr;
    c2<double>* Jm = nullptr;
    c2<double>* Jn = nullptr;
    int* offs = nullptr;
    DevRule r4, r10, r20;
    int st = OK;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        if (G) (void)hipFree(G);
        if (Jm) (void)hipFree(Jm);
        if (Jn) (void)hipFree(Jn);
        if (offs) (void)hipFree(offs);
        if (full) (void)hipFree(full);
        if (r4.buf) (void)hipFree(r4.buf);
        if (r10.buf) (void)hipFree(r10.buf);
        if (r20.buf) (void)hipFree(r20.buf);
    };
#define TRY(x) do { st = (x); if (st) { cleanup(); return st; } } while (0)
#define TRYH(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); set_error(#x " failed"); cleanup(); return ERR_HIP; } } while (0)
    const int nmed = 125;
    TRYH(hipMalloc(&G, nv * 3 * sizeof(c2<double>)));
    TRYH(hipMalloc(&Jm, (size_t)nmed * 12 * 3 * sizeof(c2<double>)));
    TRYH(hipMalloc(&Jn, (size_t)8 * 21 * 3 * sizeof(c2<double>)));
    TRYH(hipMalloc(&offs, (size_t)(nmed + 8) * 3 * sizeof(int)));
    std::vector<int> ho;
    for (int k = 0; k < 5; ++k)
        for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i) { ho.push_back(i); ho.push_back(j); ho.push_back(k); }
    for (int code = 0; code < 8; ++code) { ho.push_back(code >> 2); ho.push_back((code >> 1) & 1); ho.push_back(code & 1); }
    TRYH(hipMemcpyAsync(offs, ho.data(), ho.size() * sizeof(int), hipMemcpyHostToDevice, s));
    TRY(make_rule(4, r4));
    TRY(make_rule(10, r10));
    TRY(make_rule(20, r20));
    // dx^4 / (j ko)^2 / (4 pi) (assembly_kop.m:117) is real
    const double coef = -(dx * dx * dx * dx) / (ko * ko) / (4.0 * M_PI);
    hipLaunchKernelGGL(k_kfar, dim3((unsigned)((nv + AT - 1) / AT)), dim3(AT), 0, s, G, L, M, N, dx, ko, coef, r4.r);
    hipLaunchKernelGGL(k_kpairs, dim3(nmed * 12), dim3(AT), 0, s, Jm, offs, dx, ko, r10.r, 0);
    hipLaunchKernelGGL(k_kpairs, dim3(8 * 21), dim3(AT), 0, s, Jn, offs + 3 * nmed, dx, ko, r20.r, 1);
    hipLaunchKernelGGL(k_kfix, dim3(1), dim3(192), 0, s, G, Jm, Jn, offs, nmed, L, M, N, coef);
    if (stage == 0) {
        TRYH(hipGetLastError());
        TRY(to_host(G, out, (size_t)nv * 3, s));
        cleanup();
        return OK;
    }
    hipLaunchKernelGGL(k_circulant, dim3(4096), dim3(256), 0, s, G, (c2<double>*)full, L, M, N, h->Px, h->Py, h->Pz,
                       1);
    TRYH(hipGetLastError());
    if (stage == 1) {
        TRY(to_host((const c2<double>*)full, out, (size_t)3 * h->Px * h->Py * h->Pz, s));
        cleanup();
        return OK;
    }
    TRY(fft3_forward_inplace(h, (c2<double>*)full, 3, s));
    TRYH(hipStreamSynchronize(s));
    {
        void* f = full;
        full = nullptr;          // ownership passes to the handle
        TRY(store_circ(h, 1, f));
    }
#undef TRY
#undef TRYH
    cleanup();
    return OK;
}

}  // namespace marie

using namespace marie;

extern "C" int marie_vie_assemble_N(marie_vie_op* hh, double dx, double ko, const double* singular) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    MARIE_HIP(hipSetDevice(h->device));
    return assemble_N(h, dx, ko, singular, -1, nullptr);
}

extern "C" int marie_vie_assemble_type(marie_vie_op* hh, int op, double dx, double ko, const double* singular,
                                       int type, double* out) {
    Op* h = reinterpret_cast<Op*>(hh);
    if (!h) { set_error("null handle"); return MARIE_ERR_ARG; }
    if ((op != MARIE_OP_N && op != MARIE_OP_K) || (type != MARIE_TYPE_T && type != MARIE_TYPE_C) || !out) {
        set_error("invalid arguments (op MARIE_OP_N or MARIE_OP_K, type MARIE_TYPE_T or MARIE_TYPE_C, out)");
        return MARIE_ERR_ARG;
    }
    MARIE_HIP(hipSetDevice(h->device));
    if (op == MARIE_OP_N) return assemble_N(h, dx, ko, singular, type, out);
    return assemble_K(h, dx, ko, type, out);
}

extern "C" int marie_vie_singular_terms(double dx, double ko, double* out10) {
    if (!out10 || !(dx > 0) || !std::isfinite(ko)) { set_error("invalid arguments"); return MARIE_ERR_ARG; }
    c2<double>* d = nullptr;
    MARIE_HIP(hipMalloc(&d, 5 * sizeof(c2<double>)));
    int st = singular_terms(dx, ko, d, nullptr);
    if (!st) {
        c2<double> h[5];
        if (hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("copy failed");
            st = MARIE_ERR_HIP;
        } else {
            for (int i = 0; i < 5; ++i) { out10[2 * i] = h[i].x; out10[2 * i + 1] = h[i].y; }
        }
    }
    (void)hipFree(d);
    return st;
}
