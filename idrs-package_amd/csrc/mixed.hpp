// Mixed-precision IDR(s): single-precision inner solves (values and vectors
// in FP32) inside FP64 defect correction.
#pragma once

#include "idrs.hpp"

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

}  // namespace idrs
