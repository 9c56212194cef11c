// P^T w for shadow spaces wider than the one-pass kernels hold (PDOT_MAX
// rows), shared by the IDR(s) and Krylov solvers.
#pragma once

#include "idrs.hpp"

namespace idrs {

// P^T w for shadow spaces wider than PDOT_MAX rows: grid row y sums the P
// components [y*PDOT_MAX, min(prow, (y+1)*PDOT_MAX)) with k_pdot's
// arithmetic and row order into partials [(voff + i) * nparts + block]; a
// finished solve (info >= 0) returns at once
template <class T>
__global__ __launch_bounds__(256) void k_pdot_wide(long n, const T *__restrict__ w, const T *__restrict__ P, int prow,
                                                   int conjP, T *partials, int nparts, int voff, const int *info) {
    if (info && *info >= 0) return;
    constexpr int NV = PDOT_MAX;
    __shared__ T lds[NV * 4];
    const int p0 = blockIdx.y * NV, cnt = min(NV, prow - p0);
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const T wq = w[q];
        const T *pr = P + q * prow + p0;
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < cnt) {
                T p = pr[i];
                if (conjP) p = Tr<T>::conj(p);
                e[i] = Tr<T>::fma(p, wq, e[i]);
            }
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < cnt; i++) partials[(long)(voff + p0 + i) * nparts + blockIdx.x] = e[i];
}

template <class T>
inline void launch_pdot_wide(long n, const T *w, const T *P, int prow, int conjP, T *partials, int voff,
                             const int *info, hipStream_t st) {
    const int g = vec_grid(n);
    hipLaunchKernelGGL((k_pdot_wide<T>), dim3(g, ceil_div(prow, PDOT_MAX)), dim3(256), 0, st, n, w, P, prow, conjP,
                       partials, g, voff, info);
    LAUNCH_CHECK();
}

}  // namespace idrs
