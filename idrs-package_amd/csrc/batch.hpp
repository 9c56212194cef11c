// Batched IDR(s): many independent systems, stored as one block-diagonal
// operator with a row partition, solved together.
#pragma once

#include "idrs.hpp"
#include "mixed.hpp"

#include <vector>

namespace idrs {

// Solves the nsys systems of A (A.K->part must be set; b, x concatenated
// device vectors) with IDR(s); every system runs the single-system
// recurrences with its own scalars and stops on its own tolerance test.
// stats: one entry per system.  Options not offered here: M1, U0, resvec.
template <class T>
void batch_idrs_solve(const Operator &A, int s, const T *b, T *x, const SolveOpts<T> &o,
                      std::vector<SolveStats> &stats, double &solve_ms, hipStream_t stream);

// Mixed precision per system of a batch: r_k = b_k - A_k x_k in FP64, then
// batched IDR(s) on A d_k = r_k/|r_k| to inner_tol with the SpMVs on the
// FP32 copy of the matrix values (vectors FP64), x_k += |r_k| d_k, until
// every system meets o.tol on its true FP64 residual (at most max_outer
// replacements; o.maxit bounds each inner solve).  A system that has
// finished gets a zero right-hand side, which the inner solve finishes at
// once.  Real matrix values, no seed shift; M1, U0, omega, resvec, x0 and
// PT_raw are not offered.  stats: one per system.
void batch_mixed_solve(const Operator &A, int s, const double *b, double *x, const SolveOpts<double> &o,
                       double inner_tol, int max_outer, std::vector<MixedStats> &stats, double &solve_ms,
                       hipStream_t stream);

}  // namespace idrs
