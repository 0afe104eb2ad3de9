// qk_eigen.cpp -- Sturm bisection, implicit QL for symmetric tridiagonal
// matrices, and a Householder-based Hermitian eigensolver (host code).
//
// These replace the reference's dense helpers (src/solvers/lanczos.cpp:25-98,
// src/solvers/ftlm.cpp:61-119: tridiagonal QR with a full eigenvector matrix
// rebuilt after every Lanczos step; src/solvers/davidson.cpp:22-82: complex
// Jacobi) with a different algorithm for each job:
//  * the running ground-state energy of the Lanczos replay needs only the
//    lowest eigenvalue: bisection on the Sturm count, O(n) per halving;
//  * FTLM needs all eigenvalues and the first eigenvector components only:
//    QL with the rotations applied to one row;
//  * the final Ritz vector needs one full eigenvector matrix once;
//  * Davidson's projected matrix is reduced to tridiagonal form first.
// Any backward-stable method gives the eigenvalues to O(eps * |T|).
#include "qk_eigen.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace qk {

namespace {

// number of eigenvalues of (d, e) strictly below x: the negative pivots of
// the LDL^T factorisation of T - x I (Sylvester's law of inertia)
int sturm_count(const double* d, const double* e, int n, double x, double pivmin)
{
    int neg = 0;
    double piv = d[0] - x;
    for (int i = 0;;) {
        if (std::fabs(piv) < pivmin) piv = -pivmin;
        if (piv < 0.0) ++neg;
        if (++i == n) break;
        piv = (d[i] - x) - (e[i - 1] * e[i - 1]) / piv;
    }
    return neg;
}

} // namespace

double symtri_lowest(const double* d, const double* e, int n)
{
    if (n <= 0) return 0.0;
    if (n == 1) return d[0];
    // Gershgorin interval; the lowest eigenvalue is also <= min d (a
    // Rayleigh quotient of a unit vector)
    double glo = d[0], ghi = d[0], dmin = d[0], emax2 = 0.0;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
        glo = std::min(glo, d[i] - r);
        ghi = std::max(ghi, d[i] + r);
        dmin = std::min(dmin, d[i]);
        if (i + 1 < n) emax2 = std::max(emax2, e[i] * e[i]);
    }
    const double tnorm = std::max(std::fabs(glo), std::fabs(ghi));
    const double pivmin = DBL_MIN * std::max(1.0, emax2);
    const double pad = 4.0 * DBL_EPSILON * tnorm + 2.0 * pivmin;
    double lo = glo - pad, hi = dmin + pad;
    const double width = 2.0 * DBL_EPSILON * tnorm;
    for (int it = 0; it < 256 && hi - lo > width; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(d, e, n, mid, pivmin) >= 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

void symtri_eig(const double* d0, const double* e0, int n, double* w, double* z, int want)
{
    if (n <= 0) return;
    std::vector<double> d(d0, d0 + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = e0[i];
    if (want == kAllVectors) {
        std::fill(z, z + (size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    } else if (want == kFirstRow) {
        std::fill(z, z + n, 0.0);
        z[0] = 1.0;
    }
    // rotation G in the plane (i, i+1): u_i <- c u_i + s u_{i+1},
    // u_{i+1} <- -s u_i + c u_{i+1}; T <- G T G^T, eigenvectors V <- V G^T
    auto rotate_vectors = [&](int i, double c, double s) {
        if (want == kAllVectors) {
            for (int r = 0; r < n; ++r) {
                double* row = z + (size_t)r * n;
                const double zi = row[i], zj = row[i + 1];
                row[i] = c * zi + s * zj;
                row[i + 1] = c * zj - s * zi;
            }
        } else if (want == kFirstRow) {
            const double zi = z[i], zj = z[i + 1];
            z[i] = c * zi + s * zj;
            z[i + 1] = c * zj - s * zi;
        }
    };
    // eigenvalues converge at the top of each unreduced block, one index at
    // a time from the first
    for (int top = 0; top < n; ++top) {
        for (int sweeps = 0;; ++sweeps) {
            int bot = top;
            for (; bot + 1 < n; ++bot) {
                if (std::fabs(e[bot]) <= DBL_EPSILON * (std::fabs(d[bot]) + std::fabs(d[bot + 1]))) {
                    e[bot] = 0.0;
                    break;
                }
            }
            if (bot == top) break;
            if (sweeps >= 60) throw std::runtime_error("tridiagonal eigensolver: no convergence");
            // shift: the eigenvalue of the leading 2x2 nearer to d[top]
            const double half = 0.5 * (d[top + 1] - d[top]);
            const double rad = std::hypot(half, e[top]);
            const double shift = d[top] - e[top] * e[top] / (half >= 0.0 ? half + rad : half - rad);
            // the first rotation maps the last column of T - shift I onto the
            // last unit vector's direction; later ones push the bulge at
            // (i, i + 2) up until it leaves the block
            double X = e[bot - 1], Y = d[bot] - shift, bulge = 0.0;
            for (int i = bot - 1; i >= top; --i) {
                if (i < bot - 1) {
                    X = bulge;
                    Y = e[i + 1];
                }
                const double r = std::hypot(X, Y);
                const double c = r > 0.0 ? Y / r : 1.0;
                const double s = r > 0.0 ? -X / r : 0.0;
                if (i < bot - 1) e[i + 1] = r;
                const double a = d[i], b = d[i + 1], f = e[i];
                const double cc = c * c, ss = s * s, cs2f = 2.0 * c * s * f;
                d[i] = cc * a + ss * b + cs2f;
                d[i + 1] = ss * a + cc * b - cs2f;
                e[i] = c * s * (b - a) + (cc - ss) * f;
                if (i > top) {
                    const double g = e[i - 1];
                    e[i - 1] = c * g;
                    bulge = -s * g;
                }
                rotate_vectors(i, c, s);
            }
        }
    }
    // ascending order (stable: equal eigenvalues keep their QL order)
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < n; ++i) w[i] = d[idx[i]];
    if (want == kFirstRow) {
        std::vector<double> t(z, z + n);
        for (int i = 0; i < n; ++i) z[i] = t[idx[i]];
    } else if (want == kAllVectors) {
        std::vector<double> t(z, z + (size_t)n * n);
        for (int r = 0; r < n; ++r)
            for (int i = 0; i < n; ++i) z[(size_t)r * n + i] = t[(size_t)r * n + idx[i]];
    }
}

void herm_eig(int n, const std::vector<std::complex<double>>& A0, std::vector<double>& w,
              std::vector<std::complex<double>>& V)
{
    using cd = std::complex<double>;
    w.assign(n, 0.0);
    V.assign((size_t)n * n, cd(0.0));
    if (n <= 0) return;
    std::vector<cd> A(A0);
    std::vector<cd> Q((size_t)n * n, cd(0.0));
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    auto at = [&](std::vector<cd>& M, int r, int c) -> cd& { return M[(size_t)r * n + c]; };
    // Householder reflectors P = I - 2 v v^* / |v|^2 on indices k+1..n-1
    // zero column k below its first subdiagonal entry: A <- P A P, Q <- Q P
    std::vector<cd> v(n), tmp(n);
    for (int k = 0; k + 2 < n; ++k) {
        double nx2 = 0.0;
        for (int i = k + 1; i < n; ++i) nx2 += std::norm(at(A, i, k));
        if (nx2 == 0.0) continue;
        const double nx = std::sqrt(nx2);
        const cd x0 = at(A, k + 1, k);
        const cd phase = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cd(1.0);
        const cd alpha = -phase * nx;   // the reflected column's subdiagonal entry
        for (int i = k + 1; i < n; ++i) v[i] = at(A, i, k);
        v[k + 1] -= alpha;
        double nv2 = 0.0;
        for (int i = k + 1; i < n; ++i) nv2 += std::norm(v[i]);
        if (nv2 == 0.0) continue;
        const double tau = 2.0 / nv2;
        // left: rows k+1.., all columns k..n-1
        for (int j = k; j < n; ++j) {
            cd sdot = 0.0;
            for (int i = k + 1; i < n; ++i) sdot += std::conj(v[i]) * at(A, i, j);
            sdot *= tau;
            for (int i = k + 1; i < n; ++i) at(A, i, j) -= v[i] * sdot;
        }
        // right: columns k+1.., all rows k..n-1
        for (int i = k; i < n; ++i) {
            cd sdot = 0.0;
            for (int j = k + 1; j < n; ++j) sdot += at(A, i, j) * v[j];
            sdot *= tau;
            for (int j = k + 1; j < n; ++j) at(A, i, j) -= sdot * std::conj(v[j]);
        }
        for (int i = k + 2; i < n; ++i) at(A, i, k) = at(A, k, i) = 0.0;
        at(A, k + 1, k) = alpha;
        at(A, k, k + 1) = std::conj(alpha);
        for (int r = 0; r < n; ++r) {
            cd sdot = 0.0;
            for (int j = k + 1; j < n; ++j) sdot += at(Q, r, j) * v[j];
            sdot *= tau;
            for (int j = k + 1; j < n; ++j) at(Q, r, j) -= sdot * std::conj(v[j]);
        }
    }
    // Hermitian tridiagonal -> real symmetric: T_real = D^* T D with
    // D_0 = 1, D_{i+1} = D_i t_i / |t_i| (t_i = T[i+1][i])
    std::vector<double> dd(n), ee(std::max(n - 1, 1), 0.0);
    std::vector<cd> D(n, cd(1.0));
    for (int i = 0; i < n; ++i) dd[i] = at(A, i, i).real();
    for (int i = 0; i + 1 < n; ++i) {
        const cd t = at(A, i + 1, i);
        const double at_ = std::abs(t);
        ee[i] = at_;
        D[i + 1] = at_ > 0.0 ? D[i] * (t / at_) : D[i];
    }
    std::vector<double> Z((size_t)n * n);
    symtri_eig(dd.data(), ee.data(), n, w.data(), Z.data(), kAllVectors);
    // eigenvectors of A: Q D Z
    for (int r = 0; r < n; ++r) {
        for (int i = 0; i < n; ++i) tmp[i] = at(Q, r, i) * D[i];
        for (int j = 0; j < n; ++j) {
            cd sacc = 0.0;
            for (int i = 0; i < n; ++i) sacc += tmp[i] * Z[(size_t)i * n + j];
            V[(size_t)r * n + j] = sacc;
        }
    }
}

} // namespace qk
