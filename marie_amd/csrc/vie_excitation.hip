// Incident fields of current elements (coil segments, dipoles) at the voxel
// centres: eval_DGF.m:37-106, the free-space dyadic Green's function of a
// point current element J (current times segment vector) at distance R:
//   E = chi [P (R R^T) - Q I] J,  chi = j w mu exp(-jkR) / (4 pi k^2 R^3),
//       Q = k^2 R^2 - (1 + jkR),  P = (Q - 2 (1 + jkR)) / R^2
//   H = exp(-jkR) (1 + jkR) / (4 pi R^3) (J x R)
// (R = observation - source). Used by CCL_Excitation.m:41-70 with the
// segments' midpoints and J = (P - N) Jc.
//
// One thread per observation point and port: the segments of a port are
// looped in order (the reference's accumulation order), their data staged
// through LDS in chunks.
#include "../../include/marie_vie.h"
#include "vie_handle.hpp"

#include <cmath>
#include <vector>

namespace marie {
namespace {

constexpr int DT = 256;
constexpr int DCH = 256;   // segments staged per chunk

__global__ __launch_bounds__(DT) void k_dgf(const double* __restrict__ sc, const c2<double>* __restrict__ sj,
                                            const int* __restrict__ port_start, const double* __restrict__ pts,
                                            int64_t no, double ko, double omega_mu, c2<double>* __restrict__ E,
                                            c2<double>* __restrict__ H) {
    __shared__ double cs[3][DCH];
    __shared__ c2<double> js[3][DCH];
    const int port = blockIdx.y;
    const int s0 = port_start[port], s1 = port_start[port + 1];
    const int64_t i = blockIdx.x * (int64_t)DT + threadIdx.x;
    const bool live = i < no;
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (live) {
        ox = pts[i];
        oy = pts[no + i];
        oz = pts[2 * no + i];
    }
    double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0}, hr[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    const double k2 = ko * ko;
    const double inv4pi = 1.0 / (4.0 * M_PI);
    for (int c0 = s0; c0 < s1; c0 += DCH) {
        const int nc = (s1 - c0) < DCH ? (s1 - c0) : DCH;
        __syncthreads();
        for (int t = threadIdx.x; t < nc; t += DT)
            for (int q = 0; q < 3; ++q) {
                cs[q][t] = sc[(int64_t)(c0 + t) * 3 + q];
                js[q][t] = sj[(int64_t)(c0 + t) * 3 + q];
            }
        __syncthreads();
        if (!live) continue;
        for (int t = 0; t < nc; ++t) {
            const double X = ox - cs[0][t], Y = oy - cs[1][t], Z = oz - cs[2][t];
            const double R2 = X * X + Y * Y + Z * Z;
            const double R = sqrt(R2);
            const double R3 = R * R2;
            double sn, cn;
            sincos(ko * R, &sn, &cn);
            // exp(-jkR) = cn - j sn
            // chi = j w mu exp(-jkR) / (4 pi k^2 R^3)
            const double a = omega_mu * inv4pi / (k2 * R3);
            const double chr = a * sn, chi_ = a * cn;        // j (cn - j sn) = sn + j cn
            // Pp = 1 + jkR ; Q = k^2 R^2 - Pp ; P = (Q - 2 Pp) / R^2
            const double qr = k2 * R2 - 1.0, qi = -ko * R;
            const double pr = (qr - 2.0) / R2, pi = (qi - 2.0 * ko * R) / R2;
            // Q and P scaled by chi
            const double Qr = chr * qr - chi_ * qi, Qi = chr * qi + chi_ * qr;
            const double Pr = chr * pr - chi_ * pi, Pi = chr * pi + chi_ * pr;
            const c2<double> J1 = js[0][t], J2 = js[1][t], J3 = js[2][t];
            const double d[3] = {X, Y, Z};
            const c2<double> Jv[3] = {J1, J2, J3};
            // E_a = sum_b (P d_a d_b - Q delta_ab) J_b
            for (int aa = 0; aa < 3; ++aa) {
                double sr = 0.0, si = 0.0;
                for (int b = 0; b < 3; ++b) {
                    double mr = Pr * d[aa] * d[b], mi = Pi * d[aa] * d[b];
                    if (aa == b) { mr -= Qr; mi -= Qi; }
                    sr += mr * Jv[b].x - mi * Jv[b].y;
                    si += mr * Jv[b].y + mi * Jv[b].x;
                }
                er[aa] += sr;
                ei[aa] += si;
            }
            // mult = exp(-jkR) (1 + jkR) / (4 pi R^3)
            const double mr = (cn + ko * R * sn) * inv4pi / R3, mi = (ko * R * cn - sn) * inv4pi / R3;
            const double Xr[3] = {mr * X, mr * Y, mr * Z}, Xi[3] = {mi * X, mi * Y, mi * Z};
            // H = (J x R) mult: Hx = Z J2 - Y J3, Hy = X J3 - Z J1, Hz = Y J1 - X J2
            auto cm = [](double ar, double ai, c2<double> b, double& outr, double& outi, double sgn) {
                outr += sgn * (ar * b.x - ai * b.y);
                outi += sgn * (ar * b.y + ai * b.x);
            };
            cm(Xr[2], Xi[2], J2, hr[0], hi[0], 1.0);
            cm(Xr[1], Xi[1], J3, hr[0], hi[0], -1.0);
            cm(Xr[0], Xi[0], J3, hr[1], hi[1], 1.0);
            cm(Xr[2], Xi[2], J1, hr[1], hi[1], -1.0);
            cm(Xr[1], Xi[1], J1, hr[2], hi[2], 1.0);
            cm(Xr[0], Xi[0], J2, hr[2], hi[2], -1.0);
        }
    }
    if (!live) return;
    for (int q = 0; q < 3; ++q) {
        E[((int64_t)port * 3 + q) * no + i] = c2<double>{er[q], ei[q]};
        H[((int64_t)port * 3 + q) * no + i] = c2<double>{hr[q], hi[q]};
    }
}

}  // namespace
}  // namespace marie

using namespace marie;

extern "C" int marie_vie_dgf_fields_device(const double* seg_center_dev, const void* seg_J_dev,
                                           const int* port_start_dev, int nport, const double* points_dev,
                                           int64_t npoints, double ko, double omega_mu, void* E_dev, void* H_dev,
                                           void* stream) {
    if (!seg_center_dev || !seg_J_dev || !port_start_dev || nport < 1 || nport > 65535 || !points_dev ||
        npoints < 0 || !(ko > 0.0) || !(omega_mu > 0.0) || !E_dev || !H_dev) {
        set_error("invalid dgf arguments");
        return MARIE_ERR_ARG;
    }
    if (npoints == 0) return MARIE_OK;
    const dim3 g((unsigned)((npoints + DT - 1) / DT), (unsigned)nport);
    hipLaunchKernelGGL(k_dgf, g, dim3(DT), 0, (hipStream_t)stream, seg_center_dev, (const c2<double>*)seg_J_dev,
                       port_start_dev, points_dev, npoints, ko, omega_mu, (c2<double>*)E_dev, (c2<double>*)H_dev);
    MARIE_HIP(hipGetLastError());
    MARIE_HIP(hipStreamSynchronize((hipStream_t)stream));
    return MARIE_OK;
}

extern "C" int marie_vie_dgf_fields(const double* seg_center, const double* seg_J, const int* port_start, int nport,
                                    const double* points, int64_t npoints, double ko, double omega_mu, double* E,
                                    double* H, int device) {
    if (!seg_center || !seg_J || !port_start || nport < 1 || !points || npoints < 0 || !E || !H) {
        set_error("invalid dgf arguments");
        return MARIE_ERR_ARG;
    }
    for (int p = 0; p < nport; ++p)
        if (port_start[p + 1] < port_start[p] || port_start[0] != 0) {
            set_error("port_start must be non-decreasing from 0");
            return MARIE_ERR_ARG;
        }
    const int ns = port_start[nport];
    for (int64_t t = 0; t < (int64_t)ns * 3; ++t)
        if (!std::isfinite(seg_center[t]) || !std::isfinite(seg_J[2 * t]) || !std::isfinite(seg_J[2 * t + 1])) {
            set_error("segment data have non-finite entries");
            return MARIE_ERR_NONFINITE;
        }
    MARIE_HIP(hipSetDevice(device));
    hipStream_t s = nullptr;
    MARIE_HIP(hipStreamCreate(&s));
    void *dc = nullptr, *dj = nullptr, *dp = nullptr, *dx = nullptr, *de = nullptr, *dh = nullptr;
    const size_t nout = (size_t)nport * 3 * npoints;
    int st = MARIE_OK;
    auto A = [&](void** p, size_t bytes) {
        if (st == MARIE_OK && hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            set_error("out of device memory for the incident fields");
            st = MARIE_ERR_ALLOC;
        }
    };
    A(&dc, (size_t)ns * 3 * sizeof(double));
    A(&dj, (size_t)ns * 3 * sizeof(c2<double>));
    A(&dp, (size_t)(nport + 1) * sizeof(int));
    A(&dx, (size_t)npoints * 3 * sizeof(double));
    A(&de, nout * sizeof(c2<double>));
    A(&dh, nout * sizeof(c2<double>));
    if (st == MARIE_OK) {
        if (hipMemcpy(dc, seg_center, (size_t)ns * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dj, seg_J, (size_t)ns * 3 * sizeof(c2<double>), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dp, port_start, (size_t)(nport + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dx, points, (size_t)npoints * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("upload failed");
            st = MARIE_ERR_HIP;
        }
    }
    if (st == MARIE_OK)
        st = marie_vie_dgf_fields_device((const double*)dc, dj, (const int*)dp, nport, (const double*)dx, npoints, ko,
                                         omega_mu, de, dh, s);
    if (st == MARIE_OK &&
        (hipMemcpy(E, de, nout * sizeof(c2<double>), hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(H, dh, nout * sizeof(c2<double>), hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("download failed");
        st = MARIE_ERR_HIP;
    }
    void* bufs[] = {dc, dj, dp, dx, de, dh};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    (void)hipStreamDestroy(s);
    return st;
}
