// QMRIDR, MSIDRS and MSQMRIDR on the device.
#pragma once

#include "idrs.hpp"

#include <complex>
#include <vector>

namespace idrs {

// Options of the three solvers (device pointers unless noted).
template <class T>
struct KrylovOpts {
    bool has_m1 = false;           // present(M1): v/M1 (identity without a diagonal)
    const Operator *M1 = nullptr;
    double tol = (double)1e-6f;
    bool has_maxit = false;
    int maxit = 0;
    int variant = 1;               // MSIDRS: 1 msidrs, 2 msbicgstab
    bool has_inner_s = false;
    int inner_s = 0;
    bool has_inner_tol = false;
    double inner_tol = 0.0;
    bool has_inner_it = false;
    int inner_it = 0;
    const T *x0 = nullptr;         // QMRIDR
    const T *omega = nullptr;      // HOST
    int n_omega = 0;
    double *resvec = nullptr;      // HOST, (resvec_len, n_sigma) column-major for the multishift solvers
    int resvec_len = 0;
    const T *PT_raw = nullptr;     // device
    unsigned long long p_seed = 1;
    T *colfac = nullptr;           // HOST output (MSIDRS), n_sigma
    bool want_colfac = false;
};

struct KrylovStats {
    int flag = -1;
    int iterations = 0;
    double relres = 1.0;              // max over shifts
    std::vector<double> relres_vec;   // per shift
    double solve_ms = 0.0;
};

// x(n) = QMRIDR(A, b, s, ...)            qmridr.inc
template <class T>
void qmridr_solve(const Operator &A, long n, int s, const T *b, T *x, const KrylovOpts<T> &o, KrylovStats &st,
                  hipStream_t stream);

// X(n, n_sigma) = MSIDRS(A, b, sigma, s, ...)   msidrs.inc  (sigma on the host)
template <class T>
void msidrs_solve(const Operator &A, long n, int n_sigma, const T *sigma, int s, const T *b, T *X,
                  const KrylovOpts<T> &o, KrylovStats &st, hipStream_t stream);

// X(n, n_sigma) = MSQMRIDR(A, b, sigma, s, ...) msqmridr.inc
template <class T>
void msqmridr_solve(const Operator &A, long n, int n_sigma, const T *sigma, int s, const T *b, T *X,
                    const KrylovOpts<T> &o, KrylovStats &st, hipStream_t stream);

}  // namespace idrs
