// Device state of one IDR(s) solve and its s x s scalar logic, shared by the
// single-system solver (idrs.hip) and the batched solver (batch.hip), so both
// run the same arithmetic in the same order.
//
// Restates MODULES/IDRS/idrs.inc:155-299 (scalar parts) and COMP_OM,
// MODULES/IDRS/idrs_module.f90:719-785.
#pragma once

#include "common.hpp"
#include "idrs.hpp"

#include <cfloat>

namespace idrs {

template <class T>
struct State {
    T M[SMAX * SMAX];          // M(i,j) -> M[i + SMAX*j]
    T f[SMAX], mu[SMAX], alpha[SMAX], beta[SMAX], gamma[SMAX], gs[SMAX];
    T om;
    double normb, tolb, normr, kappa, rho, tol;
    int info, iter, jj, maxit, s, method, prow, conjP, user_omega, n_omega, inispace;
    int rv_cap;                // resvec entries available on the device
    T *H;                      // Hessenberg matrix H(nritz+1, nritz), column-major, or null
    int nritz;
};

__device__ inline double sdiv(double a, double b) { return a / b; }
__device__ inline float sdiv(float a, float b) { return a / b; }
__device__ inline c128 sdiv(c128 a, c128 b) { return a / b; }
__device__ inline c64 sdiv(c64 a, c64 b) { return a / b; }
__device__ inline double sabs(double a) { return fabs(a); }
__device__ inline double sabs(float a) { return fabs((double)a); }
__device__ inline double sabs(c128 a) { return hypot(a.re, a.im); }
__device__ inline double sabs(c64 a) { return hypot((double)a.re, (double)a.im); }
__device__ inline double sre(double a) { return a; }
__device__ inline double sre(float a) { return a; }
__device__ inline double sre(c128 a) { return a.re; }
__device__ inline double sre(c64 a) { return a.re; }

template <class T>
__device__ inline T fmaT(T a, T b, T c) { return Tr<T>::fma(a, b, c); }

// Initial state (idrs.inc:30-126 scalars).
template <class T>
__device__ inline void scalar_init(State<T> *st, int s, int method, int prow, int conjP, int maxit, double kappa,
                                   double tol, int user_omega, int n_omega, int inispace, int rv_cap) {
    st->rv_cap = rv_cap;
    st->s = s;
    st->method = method;
    st->prow = prow;
    st->conjP = conjP;
    st->maxit = maxit;
    st->kappa = kappa;
    st->tol = tol;
    st->user_omega = user_omega;
    st->n_omega = n_omega;
    st->inispace = inispace;
    st->om = T(1.0);
    st->H = nullptr;
    st->nritz = 0;
    st->info = -1;
    st->iter = 0;
    st->jj = 0;
    for (int j = 0; j < s; j++)
        for (int i = 0; i < s; i++) st->M[i + SMAX * j] = T(0);
}

// ---------------------------------------------------------------------------
// s x s scalar logic (one thread); contraction off so each operation rounds
// like the reference's statements.
// ---------------------------------------------------------------------------
#define IDRS_MS(i, j) st->M[(i) + SMAX * (j)]

// gamma(i), i = k..s-1: forward substitution on the lower-triangular M
// (idrs.inc:159-165)
template <class T>
__device__ inline void scalar_gamma(State<T> *st, int k) {
#pragma clang fp contract(off)
    const int s = st->s;
    for (int i = k; i < s; i++) {
        T g = st->f[i];
        for (int j = k; j < i; j++) g = g - IDRS_MS(i, j) * st->gamma[j];
        st->gamma[i] = sdiv(g, IDRS_MS(i, i));
    }
}

// H(row, col), 1-based, of the Hessenberg output
template <class T>
__device__ inline T &hess_at(State<T> *st, int row, int col) {
    return st->H[(row - 1) + (long)(st->nritz + 1) * (col - 1)];
}

// after G(:,k) = A*U(:,k) and mu = P^T G(:,k)  (idrs.inc:198-229), with the
// Hessenberg columns of the inner step ii = jj s + k + 1 (idrs.inc:179-181,
// 211-213, 231-237)
template <class T>
__device__ inline void scalar_mu(State<T> *st, int k, const T *red) {
#pragma clang fp contract(off)
    const int s = st->s, prow = st->prow;
    const int ii = st->jj * s + k + 1;
    const bool out_h = st->H != nullptr && ii <= st->nritz;
    if (out_h && st->jj > 0)                                               // idrs.inc:180-181
        for (int i = k; i < s; i++) hess_at(st, ii - s + i - k, ii) = -sdiv(st->gamma[i], st->beta[i]);
    T mu[SMAX];
    for (int i = 0; i < s; i++) mu[i] = i < prow ? red[i] : T(0);
    for (int i = 0; i < k; i++) {
        T a = mu[i];
        for (int j = 0; j < i; j++) a = a - IDRS_MS(i, j) * st->alpha[j];
        a = sdiv(a, IDRS_MS(i, i));
        st->alpha[i] = a;
        for (int q = k; q < s; q++) mu[q] = mu[q] - IDRS_MS(q, i) * a;
    }
    for (int q = k; q < s; q++) IDRS_MS(q, k) = mu[q];
    if (out_h)                                                             // idrs.inc:212-213
        for (int i = 0; i < k; i++) hess_at(st, ii - k + i, ii) = sdiv(st->alpha[i], st->beta[i]);
    if (sabs(IDRS_MS(k, k)) <= DBL_MIN) {                                  // idrs.inc:216
        st->info = 3;
        return;
    }
    T beta = sdiv(st->f[k], IDRS_MS(k, k));                                // idrs.inc:222
    st->beta[k] = beta;
    for (int q = k + 1; q < s; q++) st->f[q] = st->f[q] - beta * IDRS_MS(q, k);   // idrs.inc:228
    if (out_h) {                                                           // idrs.inc:232-237
        hess_at(st, ii, ii) = sdiv(T(1.0), beta);
        const int l = ii - s > 1 ? ii - s : 1;
        for (int i = ii + 1; i >= l + 1; i--) hess_at(st, i, ii) = hess_at(st, i, ii) - hess_at(st, i - 1, ii);
        for (int i = l; i <= ii + 1; i++) hess_at(st, i, ii) = sdiv(hess_at(st, i, ii), st->om);
    }
    if (k + 1 < s && st->jj > 0) scalar_gamma(st, k + 1);
}

// normr, iteration count, convergence after an inner step (idrs.inc:240-249)
template <class T>
__device__ inline void scalar_norm(State<T> *st, double nr, double *resvec) {
    st->normr = nr;
    int it = st->iter + 1;
    st->iter = it;
    if (resvec && it < st->rv_cap) resvec[it] = nr;
    if (nr < st->tolb) st->info = 0;
    else if (it == st->maxit) st->info = 1;
}

// omega (idrs.inc:260-284; COMP_OM idrs_module.f90:719-785); red = (r,r), (t,t), (t,r)
template <class T>
__device__ inline void scalar_om(State<T> *st, const T *red, const T *omega) {
#pragma clang fp contract(off)
    const int jj = st->jj + 1;
    st->jj = jj;
    T om;
    if (st->user_omega) {
        int i = jj % st->n_omega;
        if (i == 0) i = st->n_omega;
        om = omega[i - 1];
    } else {
        const T rr = red[0], tt = red[1], tr = red[2];
        const double kappa = st->kappa;
        if (kappa == 0.0) {
            om = sdiv(tr, tt);
        } else {
            double rho = sabs(tr) / sqrt(sabs(rr * tt));
            om = sdiv(tr, tt);
            if (rho < kappa) {
                om = om * T(kappa);
                om = sdiv(om, T(rho));
            }
        }
    }
    st->om = om;
    if (sabs(om) <= DBL_EPSILON) st->info = 3;                             // idrs.inc:281
}

// end of the omega step (idrs.inc:292-299) and the next cycle's f = P^T r
// and gamma (idrs.inc:147, 159-165); red = ||r||^2, P^T r
template <class T>
__device__ inline void scalar_cycle(State<T> *st, const T *red, double *resvec) {
    const int prow = st->prow;
    double nr = sqrt(sre(red[0]));
    st->normr = nr;
    int it = st->iter + 1;
    st->iter = it;
    if (resvec && it < st->rv_cap) resvec[it] = nr;
    if (nr < st->tolb) {
        st->info = 0;
        return;
    } else if (it == st->maxit) {
        st->info = 1;
        return;
    }
    for (int i = 0; i < st->s; i++) st->f[i] = i < prow ? red[1 + i] : T(0);
    scalar_gamma(st, 0);
}
#undef IDRS_MS

}  // namespace idrs
