// op_out = getOPERATORS_hip(r, f, op, type, singular)
//
// src_vie/src_operators/getOPERATORS.m:1 with the assembly on the device:
// op 'N' (default) or 'K'; type 'FFTofC' (default: the FFT of the
// circulant, what mv_AN / mv_K take), 'T' (the Toeplitz generator,
// L x M x N x 6 or x 3) or 'C' (the circulant); singular 'DIRECTFN' (the
// default of getOPERATORS.m:40-42, the reference's DIRECTFN self-terms with
// Np_1D = 20, assembly_nop.m:16) or 'DEMCEM' (machine-precision self-terms),
// ignored for 'K'. dx = r(2,1,1,1) - r(1,1,1,1), ko = 2 pi f / c0
// (getOPERATORS.m:47-48). The circulant sides are 2L, 2M, 2N as
// getOPERATORS.m builds them unless a side has a prime factor above 13 (no
// FFT plan); it is then the smallest even side >= 2L whose factors are at
// most 7 -- any side >= 2L - 1 embeds the same convolution, and mv_AN_hip
// takes the sides from fN itself. Nothing is kept between calls
// (getOPERATORS_hip('clear') does nothing).
//
// Build on a MATLAB host:
//   mex -O -I<repo>/include getOPERATORS_hip.cpp -L<repo>/marie_amd -lmarie_vie -Wl,-rpath,<repo>/marie_amd
#include "marie_mex.hpp"

#include <cmath>
#include <cstdio>
#include <string>

namespace {

int largest_prime_factor(int n) {
    int f = 1;
    for (int p = 2; p * p <= n; ++p)
        while (n % p == 0) {
            f = p;
            n /= p;
        }
    return n > 1 ? (n > f ? n : f) : f;
}

int circulant_side(int n) {
    int m = 2 * n;
    if (largest_prime_factor(m) <= 13) return m;
    while (largest_prime_factor(m) > 7) m += 2;
    return m;
}

std::string arg_string(const mxArray* a, const char* id, const char* what) {
    char buf[32] = {0};
    if (!mxIsChar(a) || mxGetString(a, buf, sizeof buf) != 0) mexErrMsgIdAndTxt(id, "%s", what);
    return std::string(buf);
}

struct Op {
    marie_vie_op* h = nullptr;
    ~Op() {
        if (h) marie_vie_destroy(h);
    }
};

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    using namespace marie_mex;
    if (is_clear(nrhs, prhs)) return;      // nothing is cached between calls
    if (nrhs < 2 || nrhs > 5) mexErrMsgIdAndTxt("MARIE:hip:nargin", "2 to 5 inputs required: r, f[, op[, type[, singular]]].");
    if (nlhs > 1) mexErrMsgIdAndTxt("MARIE:hip:nargout", "1 output required.");
    const mxArray* ra = prhs[0];
    const mwSize* rd = mxGetDimensions(ra);
    if (!mxIsDouble(ra) || mxIsComplex(ra) || mxGetNumberOfDimensions(ra) != 4 || rd[3] != 3 || rd[0] < 2)
        mexErrMsgIdAndTxt("MARIE:hip:r", "r must be a real L x M x N x 3 grid with L >= 2.");
    const std::string op = nrhs > 2 ? arg_string(prhs[2], "MARIE:hip:op", "op must be 'N' or 'K'.") : "N";
    const std::string type =
        nrhs > 3 ? arg_string(prhs[3], "MARIE:hip:type", "type must be 'FFTofC', 'T' or 'C'.") : "FFTofC";
    const std::string sing =
        nrhs > 4 ? arg_string(prhs[4], "MARIE:hip:singular", "singular must be 'DIRECTFN' or 'DEMCEM'.") : "DIRECTFN";
    if (op != "N" && op != "K") mexErrMsgIdAndTxt("MARIE:hip:op", "op must be 'N' or 'K'.");
    if (type != "FFTofC" && type != "T" && type != "C")
        mexErrMsgIdAndTxt("MARIE:hip:type", "type must be 'FFTofC', 'T' or 'C'.");
    if (sing != "DIRECTFN" && sing != "DEMCEM")
        mexErrMsgIdAndTxt("MARIE:hip:singular", "singular must be 'DIRECTFN' or 'DEMCEM'.");
    const int L = (int)rd[0], M = (int)rd[1], N = (int)rd[2];
    const double* r = mxGetPr(ra);
    const double dx = r[1] - r[0];                      // r(2,1,1,1) - r(1,1,1,1)
    const double pi = 3.14159265358979323846;
    const double ko = 2.0 * pi / 299792458.0 * mxGetScalar(prhs[1]);
    const int P[3] = {circulant_side(L), circulant_side(M), circulant_side(N)};
    Op o;
    // mexErrMsgIdAndTxt does not return and may skip destructors: release the
    // handle before raising
    auto check = [&o](int st) {
        if (!st) return;
        char msg[512];
        std::snprintf(msg, sizeof msg, "%s", marie_vie_last_error());
        if (o.h) marie_vie_destroy(o.h);
        o.h = nullptr;
        mexErrMsgIdAndTxt("MARIE:hip:status", "libmarie_vie: %s", msg);
    };
    check(marie_vie_create(&o.h, L, M, N, P[0], P[1], P[2], MARIE_PREC_F64, 0));
    const bool isN = op == "N";
    const int nc = isN ? 6 : 3;
    double sv[10];
    const double* singular = nullptr;
    if (isN && sing == "DIRECTFN") {
        check(marie_vie_singular_terms_directfn(dx, ko, 20, sv));
        singular = sv;
    }
    mwSize dims[4];
    if (type == "T") {
        dims[0] = L; dims[1] = M; dims[2] = N;
    } else {
        dims[0] = P[0]; dims[1] = P[1]; dims[2] = P[2];
    }
    dims[3] = nc;
    const size_t count = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
    std::vector<double> out(2 * count);
    if (type == "FFTofC") {
        if (isN) {
            check(marie_vie_assemble_N(o.h, dx, ko, singular));
            check(marie_vie_get_fN(o.h, out.data()));
        } else {
            check(marie_vie_assemble_K(o.h, dx, ko));
            check(marie_vie_get_fK(o.h, out.data()));
        }
    } else {
        check(marie_vie_assemble_type(o.h, isN ? MARIE_OP_N : MARIE_OP_K, dx, ko, singular,
                                     type == "T" ? MARIE_TYPE_T : MARIE_TYPE_C, out.data()));
    }
    plhs[0] = mxCreateNumericArray(4, dims, mxDOUBLE_CLASS, mxCOMPLEX);
    double* re = mxGetPr(plhs[0]);
    double* im = mxGetPi(plhs[0]);
    for (size_t t = 0; t < count; ++t) {
        re[t] = out[2 * t];
        im[t] = out[2 * t + 1];
    }
}
