// Small dense complex linear algebra for the Krylov solvers' host side
// (matrices of at most a few hundred rows, column-major std::vector):
// Householder least squares and thin QR, Cholesky solves, and the
// eigenvalues / eigenvectors of a general complex matrix (Hessenberg
// reduction, single-shift complex QR with Wilkinson shifts, back substitution
// on the Schur form). Plain C++: compiled into the library and, for tests, on
// the CPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace marie {
namespace dense {

using cd = std::complex<double>;
using Mat = std::vector<cd>;   // column-major

inline double abs1(cd z) { return std::abs(z.real()) + std::abs(z.imag()); }

// Householder vector of x (length len, stride 1): on return x holds v with
// v[0] = 1 implied, beta such that (I - beta v v^H) x = alpha e1.
inline void house(cd* x, int len, cd& alpha, double& beta) {
    double sig = 0.0;
    for (int i = 1; i < len; ++i) sig += std::norm(x[i]);
    const cd x0 = x[0];
    const double nx = std::sqrt(std::norm(x0) + sig);
    if (nx == 0.0) {
        beta = 0.0;
        alpha = 0.0;
        return;
    }
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0, 0.0);
    alpha = -ph * nx;
    const cd v0 = x0 - alpha;
    for (int i = 1; i < len; ++i) x[i] /= v0;
    // beta = 2 / (v^H v) with v = [1, x1..]
    double vv = 1.0;
    for (int i = 1; i < len; ++i) vv += std::norm(x[i]);
    beta = 2.0 / vv;
    x[0] = 1.0;
}

// QR of A (m x n, m >= n) in place: R in the upper triangle, the Householder
// vectors below it (v[0] = 1 implied), betas in `beta`.
inline void qr_house(Mat& A, int m, int n, std::vector<double>& beta) {
    beta.assign(n, 0.0);
    std::vector<cd> v(m);
    for (int k = 0; k < n; ++k) {
        const int len = m - k;
        for (int i = 0; i < len; ++i) v[i] = A[(size_t)k * m + k + i];
        cd alpha;
        double b;
        house(v.data(), len, alpha, b);
        beta[k] = b;
        // apply (I - b v v^H) to columns k..n-1
        for (int j = k; j < n; ++j) {
            cd s = 0.0;
            for (int i = 0; i < len; ++i) s += std::conj(v[i]) * A[(size_t)j * m + k + i];
            s *= b;
            for (int i = 0; i < len; ++i) A[(size_t)j * m + k + i] -= s * v[i];
        }
        for (int i = 1; i < len; ++i) A[(size_t)k * m + k + i] = v[i];
    }
}

// y = Q^H y for the factor of qr_house
inline void qr_apply_qh(const Mat& A, int m, int n, const std::vector<double>& beta, cd* y) {
    for (int k = 0; k < n; ++k) {
        cd s = y[k];
        for (int i = 1; i < m - k; ++i) s += std::conj(A[(size_t)k * m + k + i]) * y[k + i];
        s *= beta[k];
        y[k] -= s;
        for (int i = 1; i < m - k; ++i) y[k + i] -= s * A[(size_t)k * m + k + i];
    }
}

// y = Q y
inline void qr_apply_q(const Mat& A, int m, int n, const std::vector<double>& beta, cd* y) {
    for (int k = n - 1; k >= 0; --k) {
        cd s = y[k];
        for (int i = 1; i < m - k; ++i) s += std::conj(A[(size_t)k * m + k + i]) * y[k + i];
        s *= beta[k];
        y[k] -= s;
        for (int i = 1; i < m - k; ++i) y[k + i] -= s * A[(size_t)k * m + k + i];
    }
}

// Least-squares solution of min ||A x - b|| (A m x n, m >= n, full column rank)
inline bool lstsq(Mat A, int m, int n, std::vector<cd> b, std::vector<cd>& x) {
    std::vector<double> beta;
    qr_house(A, m, n, beta);
    qr_apply_qh(A, m, n, beta, b.data());
    x.assign(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        cd s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[(size_t)j * m + i] * x[j];
        const cd d = A[(size_t)i * m + i];
        if (d == cd(0.0, 0.0)) return false;
        x[i] = s / d;
    }
    return true;
}

// Thin QR: Q (m x n, orthonormal columns) and R (n x n upper triangular)
inline void qr_thin(Mat A, int m, int n, Mat& Q, Mat& R) {
    std::vector<double> beta;
    qr_house(A, m, n, beta);
    R.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) R[(size_t)j * n + i] = A[(size_t)j * m + i];
    Q.assign((size_t)m * n, 0.0);
    for (int j = 0; j < n; ++j) {
        Q[(size_t)j * m + j] = 1.0;
        qr_apply_q(A, m, n, beta, &Q[(size_t)j * m]);
    }
}

// X = X R^{-1} (X: m x n, R: n x n upper triangular)
inline bool right_solve_upper(Mat& X, int m, int n, const Mat& R) {
    for (int j = 0; j < n; ++j) {
        const cd d = R[(size_t)j * n + j];
        if (d == cd(0.0, 0.0)) return false;
        for (int i = 0; i < m; ++i) {
            cd s = X[(size_t)j * m + i];
            for (int l = 0; l < j; ++l) s -= X[(size_t)l * m + i] * R[(size_t)j * n + l];
            X[(size_t)j * m + i] = s / d;
        }
    }
    return true;
}

// Solve B X = A for Hermitian positive definite B (n x n), A n x k, by Cholesky
inline bool chol_solve(Mat B, int n, Mat& A, int k) {
    for (int j = 0; j < n; ++j) {
        double d = B[(size_t)j * n + j].real();
        for (int l = 0; l < j; ++l) d -= std::norm(B[(size_t)l * n + j]);
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        B[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            cd s = B[(size_t)j * n + i];
            for (int l = 0; l < j; ++l) s -= B[(size_t)l * n + i] * std::conj(B[(size_t)l * n + j]);
            B[(size_t)j * n + i] = s / d;
        }
    }
    // B = L L^H with L in the lower triangle
    for (int c = 0; c < k; ++c) {
        cd* y = &A[(size_t)c * n];
        for (int i = 0; i < n; ++i) {
            cd s = y[i];
            for (int l = 0; l < i; ++l) s -= B[(size_t)l * n + i] * y[l];
            y[i] = s / B[(size_t)i * n + i].real();
        }
        for (int i = n - 1; i >= 0; --i) {
            cd s = y[i];
            for (int l = i + 1; l < n; ++l) s -= std::conj(B[(size_t)i * n + l]) * y[l];
            y[i] = s / B[(size_t)i * n + i].real();
        }
    }
    return true;
}

// Complex Givens rotation G = [c s; -conj(s) c] with G [a; b] = [r; 0], c real
inline void givens(cd a, cd b, double& c, cd& s) {
    const double na = std::abs(a), nb = std::abs(b);
    if (nb == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (na == 0.0) {
        c = 0.0;
        s = std::conj(b) / nb;
    } else {
        const double t = std::hypot(na, nb);
        c = na / t;
        s = (a / na) * std::conj(b) / t;
    }
}

// Eigen-decomposition of a general complex n x n matrix A: vals (n), vecs
// (n x n, columns of unit 2-norm). Returns false when the QR iteration does
// not converge.
inline bool eig(Mat A, int n, std::vector<cd>& vals, Mat& vecs) {
    vals.assign(n, 0.0);
    vecs.assign((size_t)n * n, 0.0);
    if (n == 0) return true;
    auto a = [&](int i, int j) -> cd& { return A[(size_t)j * n + i]; };
    Mat Z((size_t)n * n, 0.0);
    auto z = [&](int i, int j) -> cd& { return Z[(size_t)j * n + i]; };
    for (int i = 0; i < n; ++i) z(i, i) = 1.0;
    // Hessenberg reduction by Householder reflections, accumulated in Z
    std::vector<cd> v(n);
    for (int k = 0; k < n - 2; ++k) {
        const int len = n - k - 1;
        for (int i = 0; i < len; ++i) v[i] = a(k + 1 + i, k);
        cd alpha;
        double b;
        house(v.data(), len, alpha, b);
        if (b == 0.0) continue;
        for (int j = 0; j < n; ++j) {          // A = (I - b v v^H) A
            cd s = 0.0;
            for (int i = 0; i < len; ++i) s += std::conj(v[i]) * a(k + 1 + i, j);
            s *= b;
            for (int i = 0; i < len; ++i) a(k + 1 + i, j) -= s * v[i];
        }
        for (int i = 0; i < n; ++i) {          // A = A (I - b v v^H)
            cd s = 0.0;
            for (int l = 0; l < len; ++l) s += a(i, k + 1 + l) * v[l];
            s *= b;
            for (int l = 0; l < len; ++l) a(i, k + 1 + l) -= s * std::conj(v[l]);
        }
        for (int i = 0; i < n; ++i) {          // Z = Z (I - b v v^H)
            cd s = 0.0;
            for (int l = 0; l < len; ++l) s += z(i, k + 1 + l) * v[l];
            s *= b;
            for (int l = 0; l < len; ++l) z(i, k + 1 + l) -= s * std::conj(v[l]);
        }
        for (int i = k + 2; i < n; ++i) a(i, k) = 0.0;
    }
    // single-shift complex QR on the Hessenberg matrix, rotations accumulated in Z
    const double eps = 2.220446049250313e-16;
    int hi = n - 1, iter = 0, total = 0;
    while (hi >= 0) {
        if (hi == 0) break;
        int l = hi;
        for (; l > 0; --l) {
            const double sc = abs1(a(l - 1, l - 1)) + abs1(a(l, l));
            if (abs1(a(l, l - 1)) <= eps * (sc > 0 ? sc : 1.0)) {
                a(l, l - 1) = 0.0;
                break;
            }
        }
        if (l == hi) {          // a(hi, hi) has converged
            --hi;
            iter = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        ++iter;
        // Wilkinson shift: eigenvalue of the trailing 2x2 nearest a(hi, hi)
        cd mu;
        if (iter % 11 == 10) {
            // exceptional shift after a run without deflation
            mu = a(hi, hi) + cd(abs1(a(hi, hi - 1)), 0.0);
        } else {
            // eigenvalues s - q r / (tr +- disc) of [p q; r s], tr = (p - s)/2:
            // the one with the larger denominator is the nearer to s and the
            // more accurately computed
            const cd p = a(hi - 1, hi - 1), q = a(hi - 1, hi), r = a(hi, hi - 1), s = a(hi, hi);
            const cd tr = 0.5 * (p - s);
            const cd disc = std::sqrt(tr * tr + q * r);
            const cd den = std::abs(tr + disc) >= std::abs(tr - disc) ? tr + disc : tr - disc;
            mu = std::abs(den) > 0 ? s - q * r / den : s;
        }
        // QR step on rows/columns l..hi: chase with Givens rotations
        cd x = a(l, l) - mu, y = a(l + 1, l);
        for (int k = l; k < hi; ++k) {
            double c;
            cd s;
            givens(x, y, c, s);
            // rows k, k+1 from column (k > l ? k-1 : l) to n-1
            for (int j = (k > l ? k - 1 : l); j < n; ++j) {
                const cd t1 = a(k, j), t2 = a(k + 1, j);
                a(k, j) = c * t1 + s * t2;
                a(k + 1, j) = -std::conj(s) * t1 + c * t2;
            }
            if (k > l) a(k + 1, k - 1) = 0.0;     // the bulge, zeroed by this rotation
            // columns k, k+1 from row 0 to min(k+2, hi)
            const int rmax = std::min(k + 2, hi);
            for (int i = 0; i <= rmax; ++i) {
                const cd t1 = a(i, k), t2 = a(i, k + 1);
                a(i, k) = c * t1 + std::conj(s) * t2;
                a(i, k + 1) = -s * t1 + c * t2;
            }
            for (int i = 0; i < n; ++i) {
                const cd t1 = z(i, k), t2 = z(i, k + 1);
                z(i, k) = c * t1 + std::conj(s) * t2;
                z(i, k + 1) = -s * t1 + c * t2;
            }
            if (k + 1 < hi) {
                x = a(k + 1, k);
                y = a(k + 2, k);
            }
        }
    }
    for (int i = 0; i < n; ++i) vals[i] = a(i, i);
    // eigenvectors of the triangular Schur factor T, then back to A's basis
    double tnorm = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, abs1(a(i, j)));
    const double small = eps * (tnorm > 0 ? tnorm : 1.0);
    std::vector<cd> w(n);
    for (int k = 0; k < n; ++k) {
        std::fill(w.begin(), w.end(), cd(0.0, 0.0));
        w[k] = 1.0;
        for (int i = k - 1; i >= 0; --i) {
            cd s = 0.0;
            for (int j = i + 1; j <= k; ++j) s += a(i, j) * w[j];
            cd d = a(i, i) - vals[k];
            if (abs1(d) < small) d = small;
            w[i] = -s / d;
        }
        double nn = 0.0;
        for (int i = 0; i < n; ++i) {
            cd s = 0.0;
            for (int j = 0; j <= k; ++j) s += z(i, j) * w[j];
            vecs[(size_t)k * n + i] = s;
            nn += std::norm(s);
        }
        nn = std::sqrt(nn);
        if (nn > 0)
            for (int i = 0; i < n; ++i) vecs[(size_t)k * n + i] /= nn;
    }
    return true;
}

}  // namespace dense
}  // namespace marie
