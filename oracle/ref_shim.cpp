// C-ABI entry points around the reference's own C++ cubature sources, so that
// Python tests can call them through ctypes. This file replaces only the MEX
// gateways (which need MATLAB's mex.h); the numerical code is compiled from
// the reference tree as it lies (see oracle/Makefile). Each gateway below
// forwards its arguments exactly as the corresponding reference gateway does:
//   cubature_nop_mex.cpp:47-59 (R_faces 3x4x6 column-major -> [3][4][6]),
//   cubature_kop_mex.cpp:75 (same argument mapping),
//   directfn_ws_st_const_mex.cpp, directfn_ws_ea_const_mex.cpp:50,
//   directfn_ws_va_const_mex.cpp:51, direct_ws_st_const_mex.cpp:49,
//   direct_ws_ea_const_mex.cpp:34, direct_ws_va_const_mex.cpp:57.
// One shared object is built per reference source (they share internal symbol
// names), selected with -DSHIM_<NAME>. Test infrastructure only.
#include <complex>
using std::complex;

#if defined(SHIM_CUBATURE_NOP)
void cubature_nop(const int Np, const double w[], const double u[], const double v[],
                  const double r_m[3], const double r_n[3], const double R_faces[3][4][6],
                  const double ko, const double dx, complex<double> I_mn[6][6]);
extern "C" void ref_cubature_nop(int Np, const double* w, const double* u, const double* v,
                                 const double* r_m, const double* r_n, const double* R_faces_cm,
                                 double ko, double dx, double* out /* 6x6 complex, [i][j] interleaved */) {
    double RR[3][4][6];
    for (int k = 0; k < 6; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 3; i++) RR[i][j][k] = R_faces_cm[i + 3 * j + 12 * k];
    complex<double> I[6][6];
    cubature_nop(Np, w, u, v, r_m, r_n, RR, ko, dx, I);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            out[2 * (6 * i + j)] = I[i][j].real();
            out[2 * (6 * i + j) + 1] = I[i][j].imag();
        }
}
#endif

#if defined(SHIM_CUBATURE_KOP)
void cubature_kop(const int Np, const double w[], const double u[], const double v[], const double r_m[3],
                  const double r_n[3], const double R_faces[3][4][6], const double ko, complex<double> IK_mn[3]);
// cubature_kop_mex.cpp:75 forwards the same arguments
extern "C" void ref_cubature_kop(int Np, const double* w, const double* u, const double* v, const double* r_m,
                                 const double* r_n, const double* R_faces_cm, double ko, double* out /* 3 complex */) {
    double RR[3][4][6];
    for (int k = 0; k < 6; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 3; i++) RR[i][j][k] = R_faces_cm[i + 3 * j + 12 * k];
    complex<double> I[3];
    cubature_kop(Np, w, u, v, r_m, r_n, RR, ko, I);
    for (int c = 0; c < 3; c++) {
        out[2 * c] = I[c].real();
        out[2 * c + 1] = I[c].imag();
    }
}
#endif

#if defined(SHIM_DIRECTFN_ST)
complex<double> directfn_ws_st_const(const double r1[], const double r2[], const double r3[],
                                     const double ko, const int Np_1D);
extern "C" void ref_directfn_st(const double* r1, const double* r2, const double* r3, double ko,
                                int Np, double* out) {
    complex<double> I = directfn_ws_st_const(r1, r2, r3, ko, Np);
    out[0] = I.real(); out[1] = I.imag();
}
#endif

#if defined(SHIM_DIRECTFN_EA)
complex<double> directfn_ws_ea_const(const double rp1[], const double rp2[], const double rp3[],
                                     const double rq1[], const double rq2[], const double rq3[],
                                     const double ko, const int Np_1D);
extern "C" void ref_directfn_ea(const double* r1, const double* r2, const double* r3,
                                const double* r4, double ko, int Np, double* out) {
    complex<double> I = directfn_ws_ea_const(r1, r2, r3, r4, r2, r1, ko, Np);
    out[0] = I.real(); out[1] = I.imag();
}
#endif

#if defined(SHIM_DIRECTFN_VA)
complex<double> directfn_ws_va_const(const double rp1[], const double rp2[], const double rp3[],
                                     const double rq1[], const double rq2[], const double rq3[],
                                     const double ko, int Np_1D);
extern "C" void ref_directfn_va(const double* r1, const double* r2, const double* r3,
                                const double* r4, const double* r5, double ko, int Np, double* out) {
    complex<double> I = directfn_ws_va_const(r1, r2, r3, r4, r5, r1, ko, Np);
    out[0] = I.real(); out[1] = I.imag();
}
#endif

#if defined(SHIM_DEMCEM_ST)
complex<double> direct_ws_st_const(const double r1[], const double r2[], const double r3[],
                                   const double ko, const int Np_1D, const double z[], const double w[]);
extern "C" void ref_demcem_st(const double* r1, const double* r2, const double* r3, double ko,
                              int Np, const double* z, const double* w, double* out) {
    complex<double> I = direct_ws_st_const(r1, r2, r3, ko, Np, z, w);
    out[0] = I.real(); out[1] = I.imag();
}
#endif

#if defined(SHIM_DEMCEM_EA)
complex<double> direct_ws_ea_const(const double r1[], const double r2[], const double r3[],
                                   const double r4[], const double ko, const int N_theta,
                                   const int N_psi, const double w_theta[], const double z_theta[],
                                   const double w_psi[], const double z_psi[]);
extern "C" void ref_demcem_ea(const double* r1, const double* r2, const double* r3,
                              const double* r4, double ko, int Nt, int Np, const double* w_t,
                              const double* z_t, const double* w_p, const double* z_p, double* out) {
    complex<double> I = direct_ws_ea_const(r1, r2, r3, r4, ko, Nt, Np, w_t, z_t, w_p, z_p);
    out[0] = I.real(); out[1] = I.imag();
}
#endif

#if defined(SHIM_DEMCEM_VA)
complex<double> direct_ws_va_const(const double r1[], const double r2[], const double r3[],
                                   const double r4[], const double r5[], const double ko,
                                   const int N_theta_p, const int N_theta_q, const int N_psi,
                                   const double w_theta_p[], const double z_theta_p[],
                                   const double w_theta_q[], const double z_theta_q[],
                                   const double w_psi[], const double z_psi[]);
extern "C" void ref_demcem_va(const double* r1, const double* r2, const double* r3,
                              const double* r4, const double* r5, double ko, int Ntp, int Ntq,
                              int Nps, const double* w_tp, const double* z_tp, const double* w_tq,
                              const double* z_tq, const double* w_ps, const double* z_ps,
                              double* out) {
    complex<double> I = direct_ws_va_const(r1, r2, r3, r4, r5, ko, Ntp, Ntq, Nps, w_tp, z_tp,
                                           w_tq, z_tq, w_ps, z_ps);
    out[0] = I.real(); out[1] = I.imag();
}
#endif
