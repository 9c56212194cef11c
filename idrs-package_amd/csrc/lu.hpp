// Shift-and-invert preconditioners: REAL_PRECON / COMPLEX_PRECON(2 | 3, K,
// M, C, shift) of MODULES/PP_MATRIX/user_module.f90:1383-1459, 1531-1602 --
// the dense (2) or band (3) matrix K [+ shift C + shift^2 M | + shift C +
// shift^2 I | - shift M | - shift I], its LU factors without pivoting
// (RLU_decomp / band_LU_decomp, :1604-1701, :2068-2179) and the triangular
// solves (RLU_solve / band_LU_solve, :1703-1831, :2181-2336), on the device.
#pragma once

#include "matrix.hpp"

#include <complex>
#include <memory>

namespace idrs {

struct LuFactor {
    int which = 0;          // 2 dense, 3 band
    long n = 0;
    int hb = 0;             // half bandwidth (get_band); dense: n - 1
    long ld = 0;            // row stride: n (dense) or 2 hb + 1 (band)
    bool cplx = false;      // factors in c128 (else double)
    // row-major factors: entry (i, c) at i*ld + c (dense) or i*ld + c - i + hb
    // (band); unit lower L below the diagonal, U on and above it
    void *lu = nullptr;
    ~LuFactor();
};

// Builds and factors; throws std::invalid_argument("Pivot equal to zero in
// LU decomposition") where the reference stops.  shift: nullptr for none.
// cplx: complex factors (COMPLEX_PRECON), else real (REAL_PRECON; every
// matrix and the shift must then be real).
std::shared_ptr<LuFactor> lu_build(int which, const DevMatrix &K, const DevMatrix *M, const DevMatrix *C,
                                   const std::complex<double> *shift, bool cplx, hipStream_t st);

// w = P^-1 v (n rows; v and w distinct)
template <class T>
void lu_solve(const LuFactor &F, const T *v, T *w, hipStream_t st);

}  // namespace idrs
