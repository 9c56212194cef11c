// Host dense linear algebra of the spectral routines (ritz.hpp).  Restates
// the reference's MODULES/IDRS/dense_la_module.f90 (QR_factorization,
// QR_solve, QR_eig, SI_power), MODULES/PP_MATRIX/matrix_module.f90
// (filter_ritzval, analyse_spectrum) and the omega ordering of
// MODULES/IDRS/ritz_module.f90 (IDRS_RITZ), statement by statement where the
// order of the arithmetic matters.
#include "ritz.hpp"

#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <type_traits>

namespace idrs {
namespace ritz {

namespace {

inline double absx(double v) { return std::fabs(v); }
inline double absx(const cd &v) { return std::abs(v); }
inline double conjx(double v) { return v; }
inline cd conjx(const cd &v) { return std::conj(v); }

// norm2(abs(q)) of a column: sqrt of the sum of squares (Fortran norm2)
template <class V>
double colnorm(const V *q, int m) {
    double s = 0.0;
    for (int i = 0; i < m; i++) s += absx(q[i]) * absx(q[i]);
    return std::sqrt(s);
}

template <class V>
int qr_fact(int m, int n, const std::vector<V> &A, std::vector<V> &Q, std::vector<V> &R) {
    Q = A;
    R.assign((size_t)n * n, V(0));
    for (int j = 0; j < n; j++) {
        V *qj = &Q[(size_t)m * j];
        for (int k = 0; k < 2; k++)          // refine = 2
            for (int i = 0; i < j; i++) {
                const V *qi = &Q[(size_t)m * i];
                V t = V(0);
                for (int l = 0; l < m; l++) t += conjx(qi[l]) * qj[l];   // dot_product(Q(:,i), Q(:,j))
                for (int l = 0; l < m; l++) qj[l] -= t * qi[l];
                R[i + (size_t)n * j] += t;
            }
        const double r = colnorm(qj, m);
        R[j + (size_t)n * j] += r;
        if (!(r > std::numeric_limits<double>::min())) return j + 1;
        for (int l = 0; l < m; l++) qj[l] /= R[j + (size_t)n * j];
    }
    return 0;
}

// X = R^-1 Q^H B (QR_solve), B m x s
template <class V>
std::vector<V> qr_solve(int m, int n, const std::vector<V> &Q, const std::vector<V> &R, const std::vector<V> &B,
                        int s) {
    std::vector<V> X((size_t)n * s, V(0));
    for (int c = 0; c < s; c++)
        for (int i = 0; i < n; i++) {
            V t = V(0);
            for (int l = 0; l < m; l++) t += conjx(Q[l + (size_t)m * i]) * B[l + (size_t)m * c];
            X[i + (size_t)n * c] = t;
        }
    for (int j = n - 1; j >= 0; j--)
        for (int c = 0; c < s; c++) {
            X[j + (size_t)n * c] /= R[j + (size_t)n * j];
            for (int i = 0; i < j; i++) X[i + (size_t)n * c] -= R[i + (size_t)n * j] * X[j + (size_t)n * c];
        }
    return X;
}

// the shifted QR iteration shared by both QR_eig variants; shift_rule 0
// (real H): the largest diagonal entry, times (1 + i) at the start; 1
// (complex H): the last diagonal entry of the active block
std::vector<cd> qr_iterate(int n, std::vector<cd> V, int maxit, int shift_rule, std::vector<cd> *U, int *converged) {
    const double tol = DBL_EPSILON;
    auto at = [&](int i, int j) -> cd & { return V[i + (size_t)n * j]; };
    if (U) {
        U->assign((size_t)n * n, cd(0));
        for (int i = 0; i < n; i++) (*U)[i + (size_t)n * i] = 1.0;
    }
    int m = n, iter = 0;
    auto largest_diag = [&](int mm) {
        cd sh = 0.0;
        for (int i = 0; i < mm; i++)
            if (std::abs(at(i, i)) > std::abs(sh)) sh = at(i, i);
        return sh;
    };
    cd shift = shift_rule == 0 ? largest_diag(m) * cd(1.0, 1.0) : at(n - 1, n - 1);
    while (m > 1 && iter < maxit) {
        for (int i = 0; i < m; i++) at(i, i) -= shift;
        std::vector<cd> Vm((size_t)m * m), Q, R;
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) Vm[i + (size_t)m * j] = at(i, j);
        if (qr_fact(m, m, Vm, Q, R) > 0) break;
        for (int j = 0; j < m; j++)         // V(1:m,1:m) = matmul(R, Q)
            for (int i = 0; i < m; i++) {
                cd t = 0.0;
                for (int l = 0; l < m; l++) t += R[i + (size_t)m * l] * Q[l + (size_t)m * j];
                at(i, j) = t;
            }
        for (int i = 0; i < m; i++) at(i, i) += shift;
        if (U) {                             // U = matmul(U, diag(Q, I))
            std::vector<cd> Un(*U);
            for (int j = 0; j < m; j++)
                for (int i = 0; i < n; i++) {
                    cd t = 0.0;
                    for (int l = 0; l < m; l++) t += (*U)[i + (size_t)n * l] * Q[l + (size_t)m * j];
                    Un[i + (size_t)n * j] = t;
                }
            *U = Un;
        }
        if (std::abs(at(m - 1, m - 2)) < tol * std::abs(at(m - 1, m - 1))) m = m - 1;
        shift = shift_rule == 0 ? largest_diag(m) : at(m - 1, m - 1);
        iter++;
    }
    std::vector<cd> eig(n);
    for (int i = 0; i < n; i++) eig[i] = at(i, i);
    if (converged) *converged = n - m + 1;
    return eig;
}

template <class V>
std::vector<V> si_power_impl(int n, const V *H, int ldh, int s, V shift, uint64_t seed, uint64_t offset) {
    std::vector<V> A((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) A[i + (size_t)n * j] = H[i + (size_t)ldh * j];
    for (int i = 0; i < n; i++) A[i + (size_t)n * i] -= shift;
    std::vector<V> Q, R, Qs, Rs;
    qr_fact(n, n, A, Q, R);
    // random start (the reference's random_number: uniform [0,1))
    std::vector<V> S((size_t)n * s);
    for (size_t q = 0; q < S.size(); q++) {
        if constexpr (std::is_same<V, cd>::value)
            S[q] = cd(uniform01(seed, offset + q), uniform01(seed, offset + S.size() + q));
        else
            S[q] = uniform01(seed, offset + q);
    }
    qr_fact(n, s, S, Qs, Rs);
    auto trace = [&]() {
        double t = 0.0;
        for (int i = 0; i < s; i++) t += std::real(V(Rs[i + (size_t)s * i]));
        return t;
    };
    const double trace0 = trace();
    double trace_old = trace0, err = 1.0;
    for (int iter = 0; err > 1e-3 && iter < 1000; iter++) {
        S = qr_solve(n, n, Q, R, Qs, s);
        qr_fact(n, s, S, Qs, Rs);
        const double t = trace();
        err = std::fabs(t - trace_old) / trace0;
        trace_old = t;
    }
    return Qs;
}

}  // namespace

int qr_factorization(int m, int n, const std::vector<cd> &V, std::vector<cd> &Q, std::vector<cd> &R) {
    return qr_fact(m, n, V, Q, R);
}
int qr_factorization(int m, int n, const std::vector<double> &V, std::vector<double> &Q, std::vector<double> &R) {
    return qr_fact(m, n, V, Q, R);
}

// RQR_eig (dense_la_module.f90:295-403): maxit 100 n, the largest-diagonal shift
std::vector<cd> qr_eig(int n, const double *H, int ldh, std::vector<cd> *U, int *converged) {
    std::vector<cd> V((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) V[i + (size_t)n * j] = H[i + (size_t)ldh * j];
    return qr_iterate(n, V, 100 * n, 0, U, converged);
}

// CQR_eig (dense_la_module.f90:192-293): maxit 10 n, shift V(m,m)
std::vector<cd> qr_eig(int n, const cd *H, int ldh, std::vector<cd> *U, int *converged) {
    std::vector<cd> V((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) V[i + (size_t)n * j] = H[i + (size_t)ldh * j];
    return qr_iterate(n, V, 10 * n, 1, U, converged);
}

std::vector<double> si_power(int n, const double *H, int ldh, int s, double shift, uint64_t seed, uint64_t offset) {
    return si_power_impl<double>(n, H, ldh, s, shift, seed, offset);
}
std::vector<cd> si_power(int n, const cd *H, int ldh, int s, cd shift, uint64_t seed, uint64_t offset) {
    return si_power_impl<cd>(n, H, ldh, s, shift, seed, offset);
}

std::vector<bool> filter_ritzval(const std::vector<cd> &rv) {
    const size_t n = rv.size();
    cd center = 0.0;
    for (const cd &v : rv) center += v;
    center /= (double)n;
    double sig = 0.0;
    for (const cd &v : rv) {
        const double a = std::abs(v - center);
        sig += a * a;
    }
    sig = std::sqrt(sig / (double)n);
    std::vector<bool> out(n);
    for (size_t i = 0; i < n; i++) out[i] = std::abs(rv[i] - center) > 1.5 * sig;
    return out;
}

namespace {
struct Box {
    double left, right, bottom, top;
};
Box box_of(const std::vector<cd> &rv, const std::vector<bool> &outlier) {
    Box b{std::numeric_limits<double>::max(), -std::numeric_limits<double>::max(),
          std::numeric_limits<double>::max(), -std::numeric_limits<double>::max()};
    for (size_t i = 0; i < rv.size(); i++) {
        if (outlier[i]) continue;
        b.left = std::min(b.left, rv[i].real());
        b.right = std::max(b.right, rv[i].real());
        b.bottom = std::min(b.bottom, rv[i].imag());
        b.top = std::max(b.top, rv[i].imag());
    }
    return b;
}
template <class C>
double radius_of(const std::vector<cd> &rv, const std::vector<bool> &outlier, C center) {
    double r = -std::numeric_limits<double>::max();
    for (size_t i = 0; i < rv.size(); i++)
        if (!outlier[i]) r = std::max(r, std::abs(rv[i] - cd(center)));
    return r;
}
}  // namespace

// ranalyse_spectrum (matrix_module.f90:1079-1183)
void analyse_spectrum(int n, const double *H, int ldh, double foci[2], double *center, double *shift) {
    const std::vector<cd> rv = qr_eig(n, H, ldh, nullptr, nullptr);
    const std::vector<bool> outlier = filter_ritzval(rv);
    const Box b = box_of(rv, outlier);
    const double center_box = (b.right + b.left) / 2.0;
    const double radius = radius_of(rv, outlier, center_box);
    double center_circle = center_box;
    if (std::fabs(center_box) / radius < 1) center_circle = center_box / (std::fabs(center_box) / radius);
    const double shift_circle = center_box - center_circle;
    const double x0 = (b.right - b.left) / 2.0, y0 = (b.top - b.bottom) / 2.0;
    const double a = std::sqrt(x0 * (x0 + y0)), bb = std::sqrt(y0 * (x0 + y0));
    double center_ellipse = center_box;
    if (std::fabs(center_box) / a < 1.0) center_ellipse = center_box / (std::fabs(center_box) / a);
    double f0, f1;
    if (a > bb) {
        const double c = std::sqrt(a * a - bb * bb);
        f0 = center_ellipse - c;
        f1 = center_ellipse + c;
    } else {
        const double c = 0.15 * center_ellipse;
        f0 = center_ellipse - c;
        f1 = center_ellipse + c;
    }
    if (foci) {
        foci[0] = f0;
        foci[1] = f1;
    }
    if (center) *center = center_circle;
    if (shift) *shift = shift_circle;
}

// canalyse_spectrum (matrix_module.f90:1185-1282)
void analyse_spectrum(int n, const cd *H, int ldh, cd foci[2], cd *center, cd *shift) {
    const std::vector<cd> rv = qr_eig(n, H, ldh, nullptr, nullptr);
    const std::vector<bool> outlier = filter_ritzval(rv);
    const Box b = box_of(rv, outlier);
    const cd center_box = cd(b.right + b.left, b.top + b.bottom) / 2.0;
    const double radius = radius_of(rv, outlier, center_box);
    cd center_circle = center_box;
    if (std::abs(center_box) / radius < 1) center_circle = center_box / (std::abs(center_box) / radius);
    const cd shift_circle = center_box - center_circle;
    const double x0 = (b.right - b.left) / 2.0, y0 = (b.top - b.bottom) / 2.0;
    const double a = std::sqrt(x0 * (x0 + y0)), bb = std::sqrt(y0 * (x0 + y0));
    cd center_ellipse = center_box;
    const double q = std::pow(center_box.real() / a, 2) + std::pow(center_box.imag() / bb, 2);
    if (q < 1.0) center_ellipse = center_box / std::sqrt(q);
    cd f0, f1;
    if (a > bb) {
        const double c = std::sqrt(a * a - bb * bb);
        f0 = center_ellipse - c;
        f1 = center_ellipse + c;
    } else {
        const double c = std::sqrt(bb * bb - a * a);
        f0 = center_ellipse - cd(0.0, c);
        f1 = center_ellipse + cd(0.0, c);
    }
    if (foci) {
        foci[0] = f0;
        foci[1] = f1;
    }
    if (center) *center = center_circle;
    if (shift) *shift = shift_circle;
}

// RIDRS_RITZ's omegas (ritz_module.f90:404-425)
std::vector<double> ritz_omegas(const std::vector<cd> &rv) {
    const size_t n = rv.size();
    std::vector<double> om(n), p(n, 1.0);
    for (size_t i = 0; i < n; i++) {
        om[i] = rv[i].real();
        if (std::fabs(om[i]) < DBL_EPSILON) om[i] = 1.0;
        om[i] = 1.0 / om[i];
    }
    for (size_t i = 0; i < n; i++) {
        size_t mx = i;
        for (size_t k = i + 1; k < n; k++)
            if (std::fabs(p[k]) > std::fabs(p[mx])) mx = k;
        const double m = om[mx];
        om[mx] = om[i];
        om[i] = m;
        for (size_t k = 0; k < n; k++) p[k] = p[k] * (1.0 - rv[k].real() * om[i]);
    }
    return om;
}

// CIDRS_RITZ's omegas (ritz_module.f90:266-286)
std::vector<cd> ritz_omegas_complex(const std::vector<cd> &rv) {
    const size_t n = rv.size();
    std::vector<cd> om(rv), p(n, cd(1.0));
    for (size_t i = 0; i < n; i++) {
        if (std::abs(om[i]) < DBL_EPSILON) om[i] = 1.0;
        om[i] = 1.0 / om[i];
    }
    for (size_t i = 0; i < n; i++) {
        size_t mx = i;
        for (size_t k = i + 1; k < n; k++)
            if (std::abs(p[k]) > std::abs(p[mx])) mx = k;
        const cd m = om[mx];
        om[mx] = om[i];
        om[i] = m;
        for (size_t k = 0; k < n; k++) p[k] = p[k] * (1.0 - rv[k] * om[i]);
    }
    return om;
}

}  // namespace ritz
}  // namespace idrs
