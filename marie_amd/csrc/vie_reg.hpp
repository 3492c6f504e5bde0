// Launchers of the compile-time-plan register-engine matvec kernels
// (instantiated for double in vie_mvp_reg_f64.hip, float in vie_mvp_reg_f32.hip).
#pragma once
#include "vie_handle.hpp"

namespace marie {

template <typename T> bool reg_supported(int P, int E);
// z pass codes: E (4, 8, 16) for the all-components kernel, 100 * TX + E for
// the kernel with one thread group per component (TX = 4 or 8 columns)
template <typename T> bool reg_z_supported(int P, int code);
// y passes: also the complex64 column-pair codes 200 + E
template <typename T> bool reg_y_supported(int P, int code);
template <typename T>
void reg_fwd_x(int P, int E, hipStream_t s, const c2<T>* x, c2<T>* A, const int* pos, const c2<T>* Mc,
               const c2<T>* tw, int L, int M, int N, int64_t nidx, int conj_in, int ncomp);
template <typename T>
void reg_strided(bool inv, int P, int E, hipStream_t s, const c2<T>* src, c2<T>* dst, const c2<T>* tw, int Px,
                 int64_t s_ls, int64_t s_plane, int64_t d_ls, int64_t d_plane, int nin, int nout, int64_t s_comp,
                 int64_t d_comp, int nplanes, int ncomp);
template <typename T>
void reg_z(bool oct, int prod, int P, int E, hipStream_t s, c2<T>* B, const c2<T>* fN, const c2<T>* tw, int Px,
           int Py, int N, int nrhs);
template <typename T>
void reg_inv_x(int P, int E, hipStream_t s, const c2<T>* A, c2<T>* y, const c2<T>* x, const int* pos, const c2<T>* Mr,
               const c2<T>* Mc, const c2<T>* tw, int L, int M, int N, int64_t nidx, int op, int transp, T Gram,
               T scale, int ncomp);

}  // namespace marie
