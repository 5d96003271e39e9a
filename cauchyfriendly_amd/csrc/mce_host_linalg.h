// Host-side dense linear algebra of the estimator's window re-initialisation.
//
// sym_eig: eigen-decomposition of a real symmetric matrix by Householder
// reduction to tridiagonal form followed by the QL algorithm with implicit
// shifts (the tred2 / tqli pair of Press et al., "Numerical Recipes in C",
// 2nd ed., sections 11.2-11.3, after the EISPACK routines TRED2 and TQL2).  The
// reference uses the same published algorithm (include/eig_solve.hpp:254-428);
// the operation order below is the published one, which is what makes the
// eigenvectors, and hence the re-initialised terms, identical to the reference's.
//
// speyers_window_init: the initial characteristic-function parameters of a window
// restarted about another window's estimate (include/cauchy_util.hpp:23-113).
#pragma once
#include <cmath>
#include <cstdio>
#include <vector>

namespace mce {
namespace host {

// sqrt(a^2 + b^2) without destructive underflow or overflow.
inline double hypot_scaled(double a, double b) {
    const double aa = fabs(a), ab = fabs(b);
    if (aa > ab) {
        const double r = ab / aa;
        return aa * sqrt(1.0 + r * r);
    }
    if (ab == 0.0) return 0.0;
    const double r = aa / ab;
    return ab * sqrt(1.0 + r * r);
}

inline double copysign_nr(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// Householder reduction of the symmetric n x n matrix z (row-major, overwritten by
// the accumulated orthogonal transform) to tridiagonal form: diagonal d, sub-diagonal
// e (e[0] = 0).
inline void householder_tridiagonal(double* z, int n, double* d, double* e) {
    auto Z = [&](int r, int c) -> double& { return z[r * n + c]; };
    for (int i = n - 1; i > 0; i--) {
        const int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; k++) scale += fabs(Z(i, k));
            if (scale == 0.0) {
                e[i] = Z(i, l);
            } else {
                for (int k = 0; k <= l; k++) {
                    Z(i, k) /= scale;
                    h += Z(i, k) * Z(i, k);
                }
                double f = Z(i, l);
                double g = (f >= 0.0) ? -sqrt(h) : sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                Z(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; j++) {
                    Z(j, i) = Z(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; k++) g += Z(j, k) * Z(i, k);
                    for (int k = j + 1; k <= l; k++) g += Z(k, j) * Z(i, k);
                    e[j] = g / h;
                    f += e[j] * Z(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; j++) {
                    f = Z(i, j);
                    g = e[j] - hh * f;
                    e[j] = g;
                    for (int k = 0; k <= j; k++) Z(j, k) -= (f * e[k] + g * Z(i, k));
                }
            }
        } else {
            e[i] = Z(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    // accumulate the transformations
    for (int i = 0; i < n; i++) {
        if (d[i] != 0.0) {
            for (int j = 0; j < i; j++) {
                double g = 0.0;
                for (int k = 0; k < i; k++) g += Z(i, k) * Z(k, j);
                for (int k = 0; k < i; k++) Z(k, j) -= g * Z(k, i);
            }
        }
        d[i] = Z(i, i);
        Z(i, i) = 1.0;
        for (int j = 0; j < i; j++) {
            Z(i, j) = 0.0;
            Z(j, i) = 0.0;
        }
    }
}

// QL iterations with implicit shifts on the tridiagonal (d, e); rotations are
// applied to the columns of z, which end up holding the eigenvectors.
inline void ql_implicit(double* d, double* e, int n, double* z) {
    auto Z = [&](int r, int c) -> double& { return z[r * n + c]; };
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; m++) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m != l) {
                if (iter++ == 30) {
                    fprintf(stderr, "[cauchy_mi355x] symmetric eigen-solver: too many QL iterations\n");
                    break;
                }
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot_scaled(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + copysign_nr(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = hypot_scaled(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        f = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * f;
                        Z(k, i) = c * Z(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

// Eigenvalues evals[i] and eigenvectors (column i of evecs, row-major n x n) of the
// symmetric matrix A.
inline void sym_eig(const double* A, double* evals, double* evecs, int n) {
    std::vector<double> e(n, 0.0);
    for (int i = 0; i < n * n; i++) evecs[i] = A[i];
    for (int i = 0; i < n; i++) evals[i] = 0.0;
    householder_tridiagonal(evecs, n, evals, e.data());
    ql_implicit(evals, e.data(), n, evecs);
}

// Initial terms (A0 rows, p0, b0) of a window restarted about the estimate (x1_hat,
// Var) after the scalar measurement z_1 = H x + v, v ~ Cauchy(0, gamma).  Var is
// boosted in place when it has an eigenvalue below cov_eig_tol.  Follows
// speyers_window_init, include/cauchy_util.hpp:23-113.
inline void speyers_window_init(int N, const double* x1_hat, double* Var, const double* H, double gamma, double z_1,
                                double* A_0, double* p_0, double* b_0, double cov_eig_tol) {
    std::vector<double> work(N * N), work2(N * N), eigs(N, 0.0);
    auto matmul = [&](const double* X, const double* Y, double* C) {  // C = X Y, N x N
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                double sum = 0.0;
                for (int k = 0; k < N; k++) sum += X[i * N + k] * Y[j + k * N];
                C[i * N + j] = sum;
            }
    };
    auto matvec = [&](const double* X, const double* v, double* y) {  // y = X v
        for (int i = 0; i < N; i++) {
            double sum = 0.0;
            for (int j = 0; j < N; j++) sum += X[i * N + j] * v[j];
            y[i] = sum;
        }
    };
    auto dot = [&](const double* x, const double* y) {
        double sum = 0.0;
        for (int i = 0; i < N; i++) sum += x[i] * y[i];
        return sum;
    };
    // eigenvalues of Var; boost the diagonal when one is below the tolerance
    sym_eig(Var, work2.data(), work.data(), N);
    double min_eig = work2[0];
    for (int i = 1; i < N; i++) min_eig = work2[i] < min_eig ? work2[i] : min_eig;
    bool boost = false;
    for (int i = 0; i < N; i++) {
        if (work2[i] < cov_eig_tol) boost = true;
        if (work2[i] < min_eig) min_eig = work2[i];
    }
    if (boost) {
        double me = work2[0];
        for (int i = 1; i < N; i++) me = work2[i] < me ? work2[i] : me;
        const double add = cov_eig_tol - me;
        fprintf(stderr, "[cauchy_mi355x] window restart: covariance has an eigenvalue below %.3e; diagonal boosted\n",
                cov_eig_tol);
        for (int i = 0; i < N; i++) Var[i * N + i] += add;
    }
    // A_0 = Var + Var H^T H Var / (gamma^2 + (z_1 - H x1_hat)^2)
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) work[i * N + j] = 0.0 + H[i] * H[j];
    matmul(Var, work.data(), work2.data());
    matmul(work2.data(), Var, work.data());
    const double Hx1 = dot(H, x1_hat);
    const double scale = gamma * gamma + (z_1 - Hx1) * (z_1 - Hx1);
    const double inv_scale = 1.0 / scale;
    for (int i = 0; i < N * N; i++) work[i] *= inv_scale;
    for (int i = 0; i < N * N; i++) A_0[i] = Var[i] + work[i] * 1.0;
    // eigenvectors of that matrix become the rows of A_0
    for (int i = 0; i < N * N; i++) work[i] = A_0[i];
    sym_eig(work.data(), eigs.data(), A_0, N);
    // b_0 = x1_hat - Var H^T (z_1 - H x1_hat) / scale
    const double scale2 = (z_1 - Hx1) / scale;
    matvec(Var, H, work.data());
    for (int i = 0; i < N; i++) work[i] *= scale2;
    for (int i = 0; i < N; i++) b_0[i] = x1_hat[i] - work[i];
    // p_0 from the eigenvalues and the eigenvectors' components along H
    for (int i = 0; i < N; i++) {  // work2 = A_0^T H
        double sum = 0.0;
        for (int j = 0; j < N; j++) sum += A_0[i + j * N] * H[j];
        work2[i] = sum;
    }
    matvec(Var, H, work.data());
    const double HVarH = dot(H, work.data());
    const double scale3 = (scale + HVarH) / gamma;
    for (int i = 0; i < N; i++) {
        double v = eigs[i] / scale3;
        v *= work2[i];
        const double sg = (double)((work2[i] > 0) - (work2[i] < 0));
        v /= sg;
        p_0[i] = v;
    }
    // transpose in place: eigenvectors as rows
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) {
            const double t = A_0[i * N + j];
            A_0[i * N + j] = A_0[j * N + i];
            A_0[j * N + i] = t;
        }
}

}  // namespace host
}  // namespace mce
