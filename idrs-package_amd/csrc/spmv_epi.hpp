// Device helpers shared by the SpMV kernels of the single-system and the
// batched solvers: value/vector arithmetic, the SpMV argument block, the
// row finish and the fused epilogues (P^T y, omega dots, norms).
#pragma once

#include "matrix.hpp"

namespace idrs {

// ---------------------------------------------------------------------------
// elementwise helpers for the value/vector type combinations
// ---------------------------------------------------------------------------
__device__ inline double divv(double x, double d) { return x / d; }
__device__ inline double divv(double x, float d) { return x / (double)d; }
__device__ inline float divv(float x, float d) { return x / d; }
template <class R>
__device__ inline cplx<R> divv(cplx<R> x, R d) { return {x.re / d, x.im / d}; }
template <class R>
__device__ inline cplx<R> divv(cplx<R> x, cplx<R> d) { return x / d; }

// ---------------------------------------------------------------------------
// SpMV epilogues
//
// The operands of a row's finish and epilogue (z, the P row, r) do not depend
// on the row sum, so the kernels load them next to the matrix entries and
// without a branch around the load: an absent z is replaced by a valid
// address whose value is ignored, and P components beyond prow re-read the
// last one and are multiplied by zero.  One memory latency then covers all
// loads of a row instead of one latency per operand.
// ---------------------------------------------------------------------------
template <class VT, class T>
struct SpmvArgs {
    int nrows, ncols;
    const int *rp, *col;
    const VT *val;
    int ndiag;
    const int *offsets;
    int off8[8];       // DIA with ndiag <= 8: the offsets by value (kernel argument registers)
    int ud;            // DIA: index of an implicit unit diagonal (not in val), INT_MAX if none
    const long *sl_ptr;   // CRS: sliced copy (val in sliced order), nullptr: CRS order
    const int *sl_col;
    int pair32;           // sliced FP32 copy: val holds (value, column) pairs, 8 bytes per entry
    const T *x;
    T xs;
    const T *z;
    T a;
    int has_seed;      // seed term: y -= seed * x/sdiag (sdiag nullptr: x)
    const VT *sdiag;
    T seed;
    T *y;
};

// y(row) = xs*(acc - seed*x(row)/sdiag(row)) + a*z(row) for the vectors x
// and z (z nullptr: no z term); zq holds z(row)
template <class VT, class T>
__device__ inline T finish_col(const SpmvArgs<VT, T> &A, const T *x, long row, T acc, const T *z, T zq, T a) {
    if (A.has_seed) acc = acc - A.seed * (A.sdiag ? divv(x[row], A.sdiag[row]) : x[row]);
    T yv = A.xs * acc;
    if (z) yv = Tr<T>::fma(a, zq, yv);
    return yv;
}

// the finish of a product of A.x (Horner vector A.z with coefficient A.a)
template <class VT, class T>
__device__ inline T finish_value(const SpmvArgs<VT, T> &A, long row, T acc, T zq) {
    return finish_col(A, A.x, row, acc, A.z, zq, A.a);
}

// Block products: the columns j < nc of x, z and y (column stride ld), with
// coefficient a[j] of z's column j.  One launch holds up to SPMV_COLS columns.
template <class T>
struct ColArgs {
    int nc;
    long ld;
    const T *x;
    const T *z;
    T *y;
    T a[SPMV_COLS];
};

// Operands of one row's epilogue: PC components of its P row (PDOT: this
// lane's components c, c+step, ...), or r(row) (OMEGA).
template <class T, int EPI, int PC>
__device__ inline void epi_load(T (&pe)[PC], const EpiArgs<T> &E, long row, int c, int step) {
    if constexpr (EPI == EPI_PDOT) {
        const T *p = E.P + row * E.prow;
        if constexpr (std::is_same<T, double>::value && PC % 2 == 0) {
            // a whole, even-length P row: 16-byte loads (the row starts
            // 16-byte aligned)
            if (step == 1 && c == 0 && E.prow == PC) {
#pragma unroll
                for (int k = 0; k < PC; k += 2) {
                    const dbl2_nt v = __builtin_nontemporal_load((const dbl2_nt *)(p + k));
                    pe[k] = v.x;
                    pe[k + 1] = v.y;
                }
                return;
            }
        }
        const int last = E.prow - 1;
#pragma unroll
        for (int k = 0; k < PC; k++) pe[k] = ldnt(p + min(c + k * step, last));
    } else if constexpr (EPI == EPI_OMEGA) {
        pe[0] = E.r[row];
    }
}

// The whole P row at p (PDOT with step 1 from component 0): as epi_load.
template <class T, int PC>
__device__ inline void epi_load_row(T (&pe)[PC], const EpiArgs<T> &E, const T *p) {
    if constexpr (std::is_same<T, double>::value && PC % 2 == 0) {
        if (E.prow == PC) {
#pragma unroll
            for (int k = 0; k < PC; k += 2) {
                const dbl2_nt v = __builtin_nontemporal_load((const dbl2_nt *)(p + k));
                pe[k] = v.x;
                pe[k + 1] = v.y;
            }
            return;
        }
    }
    const int last = E.prow - 1;
#pragma unroll
    for (int k = 0; k < PC; k++) pe[k] = ldnt(p + min(k, last));
}

// Accumulates a row into the epilogue sums: PDOT adds P components
// c, c+step, ... (zero beyond prow) times y into e[0..PC); OMEGA adds
// (r,r), (y,y), (y,r); NORM adds |y|^2.
template <class T, int EPI, int NV, int PC>
__device__ inline void epi_accum(T (&e)[NV], const T (&pe)[PC], const EpiArgs<T> &E, T yv, int c, int step) {
    using TT = Tr<T>;
    if constexpr (EPI == EPI_PDOT) {
#pragma unroll
        for (int k = 0; k < PC; k++) {
            T pv = E.conjP ? TT::conj(pe[k]) : pe[k];
            pv = (c + k * step < E.prow) ? pv : T(0);
            e[k] = TT::fma(pv, yv, e[k]);
        }
    } else if constexpr (EPI == EPI_OMEGA) {
        const T rr = pe[0];
        e[0] = TT::fma(TT::conj(rr), rr, e[0]);
        e[1] = TT::fma(TT::conj(yv), yv, e[1]);
        e[2] = TT::fma(TT::conj(yv), rr, e[2]);
    } else if constexpr (EPI == EPI_NORM) {
        e[0] = e[0] + T(TT::abs2(yv));
    }
}

// Block partials of the epilogue sums where every thread holds all NV sums.
template <class T, int EPI, int NV>
__device__ inline void epi_store(T (&e)[NV], const EpiArgs<T> &E, T *lds) {
    if constexpr (EPI != EPI_NONE) {
        block_sum<T, NV>(e, lds);
        if (threadIdx.x == 0)
#pragma unroll
            for (int i = 0; i < NV; i++) E.partials[(long)i * E.nparts + blockIdx.x] = e[i];
    }
}

// Block partials of P^T y when lane g of every G-lane group holds the
// components g, g+G, ... in e[0..PC): a butterfly over the lanes of equal g
// inside the wavefront, then one LDS slot per (component, wavefront).
template <class T, int NV, int G, int PC>
__device__ inline void epi_store_group(T (&e)[PC], const EpiArgs<T> &E, T *lds) {
    constexpr int NW = SPMV_THREADS / WAVE;
#pragma unroll
    for (int k = 0; k < PC; k++)
#pragma unroll
        for (int m = G; m < WAVE; m <<= 1) e[k] = e[k] + shfl_xor(e[k], m);
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    if (lane < G)
#pragma unroll
        for (int k = 0; k < PC; k++) {
            const int i = lane + k * G;
            if (i < NV) lds[i * NW + w] = e[k];
        }
    __syncthreads();
    if (threadIdx.x < NV) {
        T sum = lds[threadIdx.x * NW];
        for (int v = 1; v < NW; v++) sum = sum + lds[threadIdx.x * NW + v];
        E.partials[(long)threadIdx.x * E.nparts + blockIdx.x] = sum;
    }
}

}  // namespace idrs
