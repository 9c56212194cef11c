/*
 * Oracle translation unit: instantiates orc_generic.inc for real (double)
 * and complex (double _Complex) arithmetic.  Test infrastructure only.
 */
#include "orc.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Default shadow-space stream: splitmix64 (S. Vigna, public domain
   generator) evaluated at a counter, top 53 bits -> [0,1).  Stands in for
   the compiler's RANDOM_NUMBER stream, which the Fortran standard leaves
   processor dependent (idrs_module.f90:851-860). */
double orc_uniform01(uint64_t seed, uint64_t index)
{
    uint64_t z = index + seed * 0xD1B54A32D192ED03ULL;
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

/* Row and vector loops run in parallel in the OpenMP build (liborc_omp.so,
   the timed CPU baseline); the checker build (liborc.so) keeps every loop
   sequential.  Parallel sums (norms, dots) then add partial sums per thread,
   a different rounding order than the sequential loops. */
#ifdef _OPENMP
#define ORC_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_FOR_SUM _Pragma("omp parallel for schedule(static) reduction(+:acc)")
#else
#define ORC_FOR
#define ORC_FOR_SUM
#endif

#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the parallel loops: sets them when n > 0, returns the count in
   use (1 in the sequential checker build). */
int orc_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ---- real instantiation ---- */
#define T double
#define FN(name) orc_r_##name
#define CONJ(x) (x)
#define ABSV(x) fabs(x)
#define RE_ONLY 1
#define ZERO 0.0
#define ONE 1.0
#include "orc_generic.inc"
#include "orc_krylov.inc"
#undef T
#undef FN
#undef CONJ
#undef ABSV
#undef RE_ONLY
#undef ZERO
#undef ONE

/* ---- complex instantiation ---- */
#define T orc_cplx
#define FN(name) orc_c_##name
#define CONJ(x) conj(x)
#define ABSV(x) cabs(x)
#define ZERO ((orc_cplx)0.0)
#define ONE ((orc_cplx)1.0)
#include "orc_generic.inc"
#include "orc_krylov.inc"

/* get_band (user_module.f90:1833-1872): the largest |column - row| of the
   stored entries (CRS, COO), the half bandwidth (CDS), or the largest
   |offset| (DIA extension), over K and, when given, C and M
   (REAL_PRECON :1423-1425) */
static int orc_band_of(const orc_user_matrix *A)
{
    int hb = 0;
    switch (A->format) {
    case ORC_CRS:
        for (int i = 0; i < A->nrows; i++)
            for (int j = A->row_ptr[i]; j <= A->row_ptr[i + 1] - 1; j++) {
                int d = abs(A->col_ind[j - 1] - (i + 1));
                if (d > hb) hb = d;
            }
        break;
    case ORC_COO:
        for (int k = 0; k < A->nnz; k++) {
            int d = abs(A->jndx[k] - A->indx[k]);
            if (d > hb) hb = d;
        }
        break;
    case ORC_CDS:
        hb = A->half_band;
        break;
    case ORC_DIA:
        for (int d = 0; d < A->ndiag; d++)
            if (abs(A->offsets[d]) > hb) hb = abs(A->offsets[d]);
        break;
    default:
        break;
    }
    return hb;
}

int orc_get_band(const orc_user_matrix *K, const orc_user_matrix *M, const orc_user_matrix *C)
{
    int hb = orc_band_of(K);
    if (C && orc_band_of(C) > hb) hb = orc_band_of(C);
    if (M && orc_band_of(M) > hb) hb = orc_band_of(M);
    return hb;
}
