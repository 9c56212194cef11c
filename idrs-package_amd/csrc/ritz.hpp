// Spectral information from the IDR(s) Hessenberg matrix (host, small dense
// matrices): the QR eigenvalue iteration, the shift-and-invert block power
// method, the spectrum analysis that turns Ritz values into polynomial
// preconditioner parameters, and the omega ordering of IDRS_RITZ.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace idrs {
namespace ritz {

using cd = std::complex<double>;

// Column-major matrix helpers: element (i, j) of an ld-strided array.
// QR_factorization (dense_la_module.f90:42-124): classical Gram-Schmidt with
// one re-orthogonalisation sweep; Q (m x n), R (n x n).  Returns 0, or the
// column j (1-based) whose norm vanished.
int qr_factorization(int m, int n, const std::vector<cd> &V, std::vector<cd> &Q, std::vector<cd> &R);
int qr_factorization(int m, int n, const std::vector<double> &V, std::vector<double> &Q, std::vector<double> &R);

// QR_eig (dense_la_module.f90:192-403) of the leading n x n block of H
// (leading dimension ldh): eigenvalues in the iteration's diagonal order;
// U (n x n Schur vectors) when non-null; *converged when non-null.
std::vector<cd> qr_eig(int n, const double *H, int ldh, std::vector<cd> *U, int *converged);
std::vector<cd> qr_eig(int n, const cd *H, int ldh, std::vector<cd> *U, int *converged);

// SI_power (dense_la_module.f90:455-603): s Schur vectors of the leading
// n x n block of H closest to shift by the block shift-and-invert power
// method, started from the uniform [0,1) stream (seed) at counter offset
// (column-major start block; complex: real parts, then imaginary parts).
std::vector<double> si_power(int n, const double *H, int ldh, int s, double shift, uint64_t seed, uint64_t offset);
std::vector<cd> si_power(int n, const cd *H, int ldh, int s, cd shift, uint64_t seed, uint64_t offset);

// filter_ritzval (matrix_module.f90:1059-1077): Ritz values farther than
// 1.5 standard deviations from their mean
std::vector<bool> filter_ritzval(const std::vector<cd> &ritzval);

// analyse_spectrum (matrix_module.f90:1079-1282) of the leading n x n
// block of H: ellipse foci (Chebyshev), circle centre (Neumann) and the
// shift that keeps the origin outside the circle
void analyse_spectrum(int n, const double *H, int ldh, double foci[2], double *center, double *shift);
void analyse_spectrum(int n, const cd *H, int ldh, cd foci[2], cd *center, cd *shift);

// omega of IDRS_RITZ (ritz_module.f90:404-425, :266-286): 1/Ritz value
// (real part in the real version; 1 where the value is below epsilon), in
// the greedy order that takes, step by step, the omega at the maximum of
// the residual polynomial built so far
std::vector<double> ritz_omegas(const std::vector<cd> &ritzval);
std::vector<cd> ritz_omegas_complex(const std::vector<cd> &ritzval);

}  // namespace ritz
}  // namespace idrs
