// Mixed-precision IDR(s): single-precision inner solves (values and vectors
// in FP32) inside FP64 defect correction.
#pragma once

#include "idrs.hpp"
#include "krylov.hpp"

namespace idrs {

struct MixedStats {
    int flag = -1;
    int iterations = 0;        // inner iterations, all outer steps
    int outer = 0;             // FP64 residual replacements
    double relres = 1.0;       // true FP64 residual
    double solve_ms = 0.0;
};

// x = A^-1 b with r = b - A x recomputed in FP64 after every inner solve
// A32 d = r/||r|| to inner_tol (SpMV on FP32 copies of the values);
// o.tol is the FP64 tolerance, o.maxit bounds the inner iterations in total.
// fp32_vectors: the inner IDR(s) keeps its vectors and scalars in FP32 too
// (T64 = double or c128, real matrix values); otherwise only the matrix
// values are FP32 and the inner solve runs in FP64 (T64 = double).
template <class T64>
void mixed_solve(const Operator &A, long n, int s, const T64 *b, T64 *x, const SolveOpts<T64> &o, double inner_tol,
                 int max_outer, bool fp32_vectors, MixedStats &st, hipStream_t stream);

struct MsMixedStats {
    int flag = -1;
    int iterations = 0;        // inner iterations: the multishift solve and every correction solve
    int outer = 0;             // FP64 residual rounds after the multishift solve
    int corrections = 0;       // single-shift correction solves
    double solve_ms = 0.0;
};

// Multishift solve with complex64 vectors (the multishift solvers' form for
// C3's complex64 + FP64-refinement variant): X(:,k) = (A - sigma_k I)^-1 b
// for every shift.  MSQMRIDR (method 0) or MSIDRS (method 1) runs on
// complex64 vectors and the FP32 copy of A's values to inner_tol; then the
// FP64 residual of every shift is recomputed, and a shift above the FP64
// tolerance o.tol gets complex64 corrections (A - sigma_k I) d = r_k/||r_k||
// by the same solver over that one shift, x_k += ||r_k|| d, until it meets
// it (at most max_outer rounds).  sigma, relres: host arrays; b, X device.
// relres: each shift's FP64 true residual.
void ms_mixed_solve(const Operator &A, long n, int ns, const c128 *sigma, int s, const c128 *b, c128 *X,
                    const KrylovOpts<c128> &o, int method, double inner_tol, int max_outer, double *relres,
                    MsMixedStats &st, hipStream_t stream);

}  // namespace idrs
