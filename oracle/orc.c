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
