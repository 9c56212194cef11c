// Batched IDR(s): many independent systems, stored as one block-diagonal
// operator with a row partition, solved together.
#pragma once

#include "idrs.hpp"

#include <vector>

namespace idrs {

// Solves the nsys systems of A (A.K->part must be set; b, x concatenated
// device vectors) with IDR(s); every system runs the single-system
// recurrences with its own scalars and stops on its own tolerance test.
// stats: one entry per system.  Options not offered here: M1, U0, resvec.
template <class T>
void batch_idrs_solve(const Operator &A, int s, const T *b, T *x, const SolveOpts<T> &o,
                      std::vector<SolveStats> &stats, double &solve_ms, hipStream_t stream);

}  // namespace idrs
