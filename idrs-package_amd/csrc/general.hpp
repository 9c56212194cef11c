// The generalized operator of matrix_module (mass and damping matrices,
// block systems, deflation): the path op_apply, op_precon and the
// scale-backs take when Operator::general() holds.  T is double or c128.
#pragma once

#include "matrix.hpp"

#include <complex>
#include <vector>

namespace idrs {

constexpr int MAX_DEFL = 16;     // deflation vectors (A%n_deflation)
constexpr int GEN_PARTS = 4096;  // most workgroups of the general path's reductions

// allocates the operator's workspace (called by op_prepare)
void gen_prepare(const Operator &op);
// w = v/A (rprecon/cprecon without shift-and-invert)
template <class T> void gen_precon(const Operator &op, const T *v, T *w, hipStream_t st);
// w = real_mv(A, v); t, u: vectors of op.rows()
template <class T> void gen_mv(const Operator &op, const T *v, T *w, T *t, T *u, hipStream_t st);
// w = Pmul(A, w)
template <class T> void gen_pmul(const Operator &op, T *w, hipStream_t st);
// w = A*v (rmatvec) with the SpMV epilogue's partials of w
template <class T>
void gen_apply(const Operator &op, const T *v, T *w, int epi, const EpiArgs<T> &E, const int *info, hipStream_t st);
// make_deflation(A, V), V device (rows x nd)
template <class T> void gen_make_deflation(Operator &op, const T *V, int nd, hipStream_t st);
// x = add_deflation(A, b, y)
template <class T> void gen_add_deflation(const Operator &op, const T *b, const T *y, T *x, hipStream_t st);
// X = ms_add_deflation(A, sigma, b, Y), columns of op.rows()
template <class T>
void gen_ms_add_deflation(const Operator &op, int nsig, const std::complex<double> *sigma, const T *b, const T *Y,
                          T *X, hipStream_t st);
// x = scaleback(A, b, y)
template <class T> void gen_scaleback(const Operator &op, const T *b, const T *y, T *x, hipStream_t st);
// one column of ms_scaleback: the shift's polynomial alpha(0:d), its
// deflation step, x/A
template <class T>
void gen_ms_scaleback_col(const Operator &op, const std::vector<std::complex<double>> &alpha, std::complex<double> sigma,
                          const T *b, const T *y, T *x, hipStream_t st);

}  // namespace idrs
