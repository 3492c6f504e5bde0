// Shared part of the MEX gateways over libmarie_vie.so (include/marie_vie.h).
//
// MATLAB conventions follow the reference's own gateways
// (src_vie/src_operators/mexcubatures/cubature_nop_mex.cpp:11-89): arity
// checks with mexErrMsgIdAndTxt, read-only inputs through mxGetPr / mxGetPi
// (separate real and imaginary storage), outputs created with
// mxCreateDoubleMatrix(..., mxCOMPLEX) and owned by MATLAB. A non-zero
// library status becomes mexErrMsgIdAndTxt("MARIE:hip:status", ...), which
// JVIE_Solver.m:326-332's try/catch turns into its CPU branch.
//
// Each gateway keeps one device operator between calls, as the reference keeps
// fN_gpu, Mr_gpu, Mc_gpu and idxS3 in the closure of JVIE_Solver.m:271-278:
// the operator is rebuilt when fN's data pointer or shape changes, the
// material is uploaded when Mr's, Mc's or Gram's identity changes, and the
// index when idx's does. Identity = data pointers, element count and a
// fingerprint of 64 sampled entries, so a Krylov loop that passes the same
// arrays every iteration uploads them once. `gateway('clear')` releases the
// operator (also done when MATLAB unloads the MEX file).
#pragma once
#include "mex.h"
#include "marie_vie.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace marie_mex {

inline void fail(int st) {
    if (st) mexErrMsgIdAndTxt("MARIE:hip:status", "libmarie_vie: %s", marie_vie_last_error());
}

// interleaved complex copy of a real or complex double array
inline void interleave(const mxArray* a, std::vector<double>& v) {
    if (!mxIsDouble(a)) mexErrMsgIdAndTxt("MARIE:hip:type", "double input expected.");
    const size_t n = mxGetNumberOfElements(a);
    v.assign(2 * n, 0.0);
    const double* re = mxGetPr(a);
    const double* im = mxIsComplex(a) ? mxGetPi(a) : nullptr;
    for (size_t i = 0; i < n; ++i) {
        v[2 * i] = re[i];
        v[2 * i + 1] = im ? im[i] : 0.0;
    }
}

struct Identity {
    const void* re = nullptr;
    const void* im = nullptr;
    size_t n = 0;
    uint64_t print = 0;
    bool operator==(const Identity& o) const { return re == o.re && im == o.im && n == o.n && print == o.print; }
};

inline uint64_t fnv(uint64_t h, const double* p, size_t n) {
    const size_t step = n > 64 ? n / 64 : 1;
    for (size_t i = 0; i < n; i += step) {
        uint64_t bits;
        std::memcpy(&bits, p + i, sizeof bits);
        h = (h ^ bits) * 1099511628211ULL;
    }
    if (n) {
        uint64_t bits;
        std::memcpy(&bits, p + n - 1, sizeof bits);
        h = (h ^ bits) * 1099511628211ULL;
    }
    return h;
}

inline Identity identity(const mxArray* a) {
    Identity id;
    id.n = mxGetNumberOfElements(a);
    id.re = mxGetPr(a);
    id.im = mxIsComplex(a) ? mxGetPi(a) : nullptr;
    uint64_t h = 1469598103934665603ULL;
    h = fnv(h, mxGetPr(a), id.n);
    if (id.im) h = fnv(h, mxGetPi(a), id.n);
    id.print = h;
    return id;
}

// The cached device operator of one gateway.
struct Cache {
    int precisions = MARIE_PREC_F64;     // MARIE_PREC_* the cached operator is created with
    marie_vie_op* op = nullptr;
    int dims[6] = {0, 0, 0, 0, 0, 0};
    Identity fN, Mr, Mc, idx;            // fN: the operator array (fN, or fK for mv_K)
    int ncomp = 0;                       // 6: fN, 3: fK
    double gram = 0.0;
    bool have_mat = false, have_idx = false, unit_mat = false;
    int64_t nidx = 0;

    void clear() {
        if (op) marie_vie_destroy(op);
        op = nullptr;
        have_mat = have_idx = unit_mat = false;
        ncomp = 0;
        fN = Mr = Mc = idx = Identity();
    }

    // fN (nc = 6) or fK (nc = 3), Px x Py x Pz x nc, on the grid of r
    // (L x M x N x 3): the operator and domain arguments of mv_N.m, mv_L.m
    // and mv_K.m (their sizes, mv_N.m:35-38)
    void operator_on_grid(const mxArray* fa, const mxArray* ra, int nc) {
        const mwSize* fd = mxGetDimensions(fa);
        const mwSize* rd = mxGetDimensions(ra);
        if (mxGetNumberOfDimensions(fa) != 4 || fd[3] != (mwSize)nc)
            mexErrMsgIdAndTxt("MARIE:hip:fN", nc == 6 ? "fN must be Px x Py x Pz x 6." : "fK must be Px x Py x Pz x 3.");
        if (mxGetNumberOfDimensions(ra) != 4 || rd[3] != 3)
            mexErrMsgIdAndTxt("MARIE:hip:r", "r must be L x M x N x 3.");
        const int d[6] = {(int)rd[0], (int)rd[1], (int)rd[2], (int)fd[0], (int)fd[1], (int)fd[2]};
        const Identity idf = identity(fa);
        if (op && idf == fN && ncomp == nc && std::memcmp(d, dims, sizeof d) == 0) return;
        clear();
        fail(marie_vie_create(&op, d[0], d[1], d[2], d[3], d[4], d[5], precisions, 0));
        std::vector<double> buf;
        interleave(fa, buf);
        fail(nc == 6 ? marie_vie_set_fN(op, buf.data()) : marie_vie_set_fK(op, buf.data()));
        fN = idf;
        ncomp = nc;
        std::memcpy(dims, d, sizeof d);
    }

    // Gram alone (mv_L.m: N x - Gram x): the material Mr = 1, Mc = 0
    void gram_only(double g) {
        if (have_mat && unit_mat && g == gram) return;
        const size_t nv = (size_t)dims[0] * dims[1] * dims[2];
        std::vector<double> one(2 * nv, 0.0), zero(2 * nv, 0.0);
        for (size_t i = 0; i < nv; ++i) one[2 * i] = 1.0;
        fail(marie_vie_set_material(op, one.data(), zero.data(), g));
        Mr = Mc = Identity();
        gram = g;
        have_mat = unit_mat = true;
    }

    // fN: Px x Py x Pz x 6; Mr: the L x M x N voxel grid
    void operator_for(const mxArray* fNa, const mxArray* Mra) {
        const mwSize* fd = mxGetDimensions(fNa);
        const mwSize* md = mxGetDimensions(Mra);
        if (mxGetNumberOfDimensions(fNa) != 4 || fd[3] != 6)
            mexErrMsgIdAndTxt("MARIE:hip:fN", "fN must be Px x Py x Pz x 6.");
        const mwSize nd = mxGetNumberOfDimensions(Mra);
        const int d[6] = {(int)md[0], nd > 1 ? (int)md[1] : 1, nd > 2 ? (int)md[2] : 1, (int)fd[0], (int)fd[1],
                          (int)fd[2]};
        const Identity idf = identity(fNa);
        if (op && idf == fN && ncomp == 6 && std::memcmp(d, dims, sizeof d) == 0) return;
        clear();
        fail(marie_vie_create(&op, d[0], d[1], d[2], d[3], d[4], d[5], precisions, 0));
        std::vector<double> buf;
        interleave(fNa, buf);
        fail(marie_vie_set_fN(op, buf.data()));
        fN = idf;
        ncomp = 6;
        std::memcpy(dims, d, sizeof d);
    }

    void material(const mxArray* Mra, const mxArray* Mca, double g) {
        const size_t nv = (size_t)dims[0] * dims[1] * dims[2];
        if (mxGetNumberOfElements(Mra) != nv || mxGetNumberOfElements(Mca) != nv)
            mexErrMsgIdAndTxt("MARIE:hip:material", "Mr and Mc must be L x M x N.");
        const Identity a = identity(Mra), b = identity(Mca);
        if (have_mat && !unit_mat && a == Mr && b == Mc && g == gram) return;
        std::vector<double> r, c;
        interleave(Mra, r);
        interleave(Mca, c);
        fail(marie_vie_set_material(op, r.data(), c.data(), g));
        Mr = a;
        Mc = b;
        gram = g;
        have_mat = true;
        unit_mat = false;
    }

    // MATLAB's 1-based linear indices into the L x M x N x 3 grid (idxS3)
    void index(const mxArray* ia) {
        if (!mxIsDouble(ia) || mxIsComplex(ia)) mexErrMsgIdAndTxt("MARIE:hip:idx", "idx must be real double.");
        const Identity a = identity(ia);
        if (have_idx && a == idx) return;
        const size_t n = mxGetNumberOfElements(ia);
        const double* p = mxGetPr(ia);
        std::vector<int64_t> v(n);
        for (size_t t = 0; t < n; ++t) v[t] = (int64_t)p[t];
        fail(marie_vie_set_index(op, v.data(), (int64_t)n, 1));
        idx = a;
        nidx = (int64_t)n;
        have_idx = true;
    }
};

// column vector of n complex values from interleaved storage
inline mxArray* complex_column(const std::vector<double>& v, size_t n) {
    mxArray* out = mxCreateDoubleMatrix((mwSize)n, 1, mxCOMPLEX);
    double* re = mxGetPr(out);
    double* im = mxGetPi(out);
    for (size_t t = 0; t < n; ++t) {
        re[t] = v[2 * t];
        im[t] = v[2 * t + 1];
    }
    return out;
}

inline bool is_clear(int nrhs, const mxArray* prhs[]) {
    if (nrhs != 1 || !mxIsChar(prhs[0])) return false;
    char s[8] = {0};
    mxGetString(prhs[0], s, sizeof s);
    return std::strcmp(s, "clear") == 0;
}

}  // namespace marie_mex
