// C entry points over the library's host-side dense complex algebra, built
// with g++ by tests/test_dense_host.py to check it against numpy on the CPU.
#include "../marie_amd/csrc/vie_dense.hpp"

using marie::dense::cd;
using marie::dense::Mat;

static Mat in(const double* p, size_t n) {
    Mat m(n);
    for (size_t i = 0; i < n; ++i) m[i] = cd(p[2 * i], p[2 * i + 1]);
    return m;
}
static void out(const Mat& m, double* p) {
    for (size_t i = 0; i < m.size(); ++i) {
        p[2 * i] = m[i].real();
        p[2 * i + 1] = m[i].imag();
    }
}

extern "C" int dh_lstsq(const double* A, int m, int n, const double* b, double* x) {
    std::vector<cd> xs;
    if (!marie::dense::lstsq(in(A, (size_t)m * n), m, n, in(b, m), xs)) return 1;
    out(xs, x);
    return 0;
}

extern "C" int dh_qr_thin(const double* A, int m, int n, double* Q, double* R) {
    Mat q, r;
    marie::dense::qr_thin(in(A, (size_t)m * n), m, n, q, r);
    out(q, Q);
    out(r, R);
    return 0;
}

extern "C" int dh_eig(const double* A, int n, double* vals, double* vecs) {
    std::vector<cd> v;
    Mat V;
    if (!marie::dense::eig(in(A, (size_t)n * n), n, v, V)) return 1;
    out(v, vals);
    out(V, vecs);
    return 0;
}

extern "C" int dh_chol_solve(const double* B, int n, double* A, int k) {
    Mat a = in(A, (size_t)n * k);
    if (!marie::dense::chol_solve(in(B, (size_t)n * n), n, a, k)) return 1;
    out(a, A);
    return 0;
}

extern "C" int dh_right_solve_upper(double* X, int m, int n, const double* R) {
    Mat x = in(X, (size_t)m * n);
    if (!marie::dense::right_solve_upper(x, m, n, in(R, (size_t)n * n))) return 1;
    out(x, X);
    return 0;
}
