// Emulation of the MEX API subset declared in tests/mex/mex.h, plus a C
// interface for ctypes: build mxArrays from host buffers, call the gateway's
// mexFunction (a MATLAB error becomes a non-zero return and a message), read
// the outputs. Test infrastructure only.
#include "mex.h"

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct mxArray_tag {
    std::vector<mwSize> dims;
    std::vector<double> re, im;
    bool cplx = false;
    bool chr = false;
    std::string str;
};

namespace {
std::vector<void (*)(void)> g_exit;
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
}  // namespace

extern "C" {
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MexError(std::string(id) + ": " + buf);
}
int mexAtExit(void (*fn)(void)) {
    for (auto f : g_exit)
        if (f == fn) return 0;
    g_exit.push_back(fn);
    return 0;
}
double* mxGetPr(const mxArray* a) { return const_cast<double*>(a->re.data()); }
double* mxGetPi(const mxArray* a) { return a->cplx ? const_cast<double*>(a->im.data()) : nullptr; }
bool mxIsComplex(const mxArray* a) { return a->cplx; }
bool mxIsDouble(const mxArray* a) { return !a->chr; }
bool mxIsChar(const mxArray* a) { return a->chr; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->chr ? a->str.size() : a->re.size(); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
double mxGetScalar(const mxArray* a) { return a->re.empty() ? 0.0 : a->re[0]; }
int mxGetString(const mxArray* a, char* buf, mwSize len) {
    if (!a->chr || len == 0) return 1;
    std::strncpy(buf, a->str.c_str(), len - 1);
    buf[len - 1] = 0;
    return a->str.size() >= len ? 1 : 0;
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
    auto* a = new mxArray_tag();
    a->dims = {m, n};
    a->re.assign(m * n, 0.0);
    a->cplx = c == mxCOMPLEX;
    if (a->cplx) a->im.assign(m * n, 0.0);
    return a;
}
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity c) {
    if (classid != mxDOUBLE_CLASS) throw MexError("emulation: double arrays only");
    auto* a = new mxArray_tag();
    size_t n = 1;
    for (mwSize d = 0; d < ndim; ++d) {
        a->dims.push_back(dims[d]);
        n *= dims[d];
    }
    a->re.assign(n, 0.0);
    a->cplx = c == mxCOMPLEX;
    if (a->cplx) a->im.assign(n, 0.0);
    return a;
}
mxArray* mxCreateDoubleScalar(double v) {
    mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
    a->re[0] = v;
    return a;
}

// ---- ctypes side
mxArray* emul_new_double(int ndims, const int64_t* dims, const double* re, const double* im) {
    auto* a = new mxArray_tag();
    size_t n = 1;
    for (int d = 0; d < ndims; ++d) {
        a->dims.push_back((mwSize)dims[d]);
        n *= (size_t)dims[d];
    }
    a->re.assign(re, re + n);
    if (im) {
        a->cplx = true;
        a->im.assign(im, im + n);
    }
    return a;
}
mxArray* emul_new_string(const char* s) {
    auto* a = new mxArray_tag();
    a->chr = true;
    a->str = s;
    a->dims = {1, (mwSize)a->str.size()};
    return a;
}
void emul_free(mxArray* a) { delete a; }
int64_t emul_numel(const mxArray* a) { return (int64_t)mxGetNumberOfElements(a); }
int emul_is_complex(const mxArray* a) { return a->cplx ? 1 : 0; }
const double* emul_re(const mxArray* a) { return a->re.data(); }
const double* emul_im(const mxArray* a) { return a->cplx ? a->im.data() : nullptr; }
int emul_dim(const mxArray* a, int d) { return d < (int)a->dims.size() ? (int)a->dims[(size_t)d] : 1; }
int emul_ndim(const mxArray* a) { return (int)a->dims.size(); }
int emul_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs, char* err, int errlen) {
    try {
        mexFunction(nlhs, plhs, nrhs, const_cast<const mxArray**>(prhs));
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
    return 0;
}
void emul_exit() {
    for (auto f : g_exit) f();
    g_exit.clear();
}
}
