// Shared device/host helpers for the IDR(s) HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cmath>

namespace idrs {

// ---------------------------------------------------------------------------
// error handling: every HIP call in host code goes through HIP_TRY, which
// records the failing call and makes the entry point return IDRS_ERR_HIP.
// ---------------------------------------------------------------------------
void set_last_error(const char *fmt, ...);

struct HipError {
    hipError_t code;
};

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::idrs::set_last_error("%s failed at %s:%d: %s", #expr, __FILE__,  \
                                   __LINE__, hipGetErrorString(e_));           \
            throw ::idrs::HipError{e_};                                        \
        }                                                                      \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

// ---------------------------------------------------------------------------
// complex<double> / complex<float> value type (plain struct: trivially
// copyable, passes as a kernel argument, 16-byte aligned loads).
// ---------------------------------------------------------------------------
template <class R>
struct alignas(2 * sizeof(R)) cplx {
    R re, im;
    cplx() = default;
    __host__ __device__ constexpr cplx(R r) : re(r), im(0) {}
    __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
};

template <class R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__host__ __device__ inline cplx<R> operator*(R a, cplx<R> b) { return {a * b.re, a * b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator*(cplx<R> b, R a) { return {a * b.re, a * b.im}; }
template <class R>
__host__ __device__ inline cplx<R> operator/(cplx<R> a, R b) { return {a.re / b, a.im / b}; }
// Smith's algorithm (R. L. Smith, CACM 5(8), 1962) for complex division.
template <class R>
__host__ __device__ inline cplx<R> operator/(cplx<R> a, cplx<R> b) {
    if (fabs(b.re) >= fabs(b.im)) {
        R r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    } else {
        R r = b.re / b.im, d = b.im + b.re * r;
        return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    }
}
template <class R>
__host__ __device__ inline bool operator==(cplx<R> a, cplx<R> b) { return a.re == b.re && a.im == b.im; }
template <class R>
__host__ __device__ inline bool operator!=(cplx<R> a, cplx<R> b) { return !(a == b); }

using c64 = cplx<float>;
using c128 = cplx<double>;

// ---------------------------------------------------------------------------
// scalar traits
// ---------------------------------------------------------------------------
__host__ __device__ inline double fmaR(double a, double b, double c) { return ::fma(a, b, c); }
__host__ __device__ inline float fmaR(float a, float b, float c) { return ::fmaf(a, b, c); }

template <class T> struct Tr;
template <> struct Tr<double> {
    using real = double;
    static constexpr bool is_complex = false;
    __host__ __device__ static double conj(double a) { return a; }
    __host__ __device__ static double abs2(double a) { return a * a; }
    __host__ __device__ static double absv(double a) { return fabs(a); }
    __host__ __device__ static double fma(double a, double b, double c) { return ::fma(a, b, c); }
    __host__ __device__ static double from_real(double a) { return a; }
};
template <> struct Tr<float> {
    using real = float;
    static constexpr bool is_complex = false;
    __host__ __device__ static float conj(float a) { return a; }
    __host__ __device__ static float abs2(float a) { return a * a; }
    __host__ __device__ static float absv(float a) { return fabsf(a); }
    __host__ __device__ static float fma(float a, float b, float c) { return ::fmaf(a, b, c); }
    __host__ __device__ static float from_real(float a) { return a; }
};
template <class R> struct Tr<cplx<R>> {
    using real = R;
    static constexpr bool is_complex = true;
    __host__ __device__ static cplx<R> conj(cplx<R> a) { return {a.re, -a.im}; }
    __host__ __device__ static R abs2(cplx<R> a) { return a.re * a.re + a.im * a.im; }
    __host__ __device__ static R absv(cplx<R> a) { return hypot(a.re, a.im); }
    // a*b + c with fused real parts
    __host__ __device__ static cplx<R> fma(cplx<R> a, cplx<R> b, cplx<R> c) {
        return {fmaR(a.re, b.re, fmaR(-a.im, b.im, c.re)), fmaR(a.re, b.im, fmaR(a.im, b.re, c.im))};
    }
    __host__ __device__ static cplx<R> from_real(R a) { return {a, R(0)}; }
};

// real matrix value times vector entry (real x complex scales both parts)
__host__ __device__ inline double mulv(double a, double x) { return a * x; }
// single-precision matrix values with double-precision vectors
__host__ __device__ inline double mulv(float a, double x) { return (double)a * x; }
__host__ __device__ inline double fmav(float a, double x, double c) { return ::fma((double)a, x, c); }
__host__ __device__ inline float mulv(float a, float x) { return a * x; }
template <class R>
__host__ __device__ inline cplx<R> mulv(R a, cplx<R> x) { return {a * x.re, a * x.im}; }
template <class R>
__host__ __device__ inline cplx<R> mulv(cplx<R> a, cplx<R> x) { return a * x; }

__host__ __device__ inline double fmav(double a, double x, double c) { return ::fma(a, x, c); }
__host__ __device__ inline float fmav(float a, float x, float c) { return ::fmaf(a, x, c); }
template <class R>
__host__ __device__ inline cplx<R> fmav(R a, cplx<R> x, cplx<R> c) { return {fmaR(a, x.re, c.re), fmaR(a, x.im, c.im)}; }
template <class R>
__host__ __device__ inline cplx<R> fmav(cplx<R> a, cplx<R> x, cplx<R> c) { return Tr<cplx<R>>::fma(a, x, c); }

// ---------------------------------------------------------------------------
// deterministic block reductions (fixed lane/wave order): wave64 butterfly,
// then the waves' partials summed in wave order by wave 0.
// ---------------------------------------------------------------------------
constexpr int WAVE = 64;

__device__ inline double shfl_xor(double v, int m) { return __shfl_xor(v, m, WAVE); }
__device__ inline float shfl_xor(float v, int m) { return __shfl_xor(v, m, WAVE); }
template <class R>
__device__ inline cplx<R> shfl_xor(cplx<R> v, int m) { return {__shfl_xor(v.re, m, WAVE), __shfl_xor(v.im, m, WAVE)}; }

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int m = WAVE / 2; m >= 1; m >>= 1) v = v + shfl_xor(v, m);
    return v;
}

// Sum NV values per thread over the block; result valid in thread 0.
// lds must hold NV * (blockDim.x / WAVE) elements.
template <class T, int NV>
__device__ inline void block_sum(T (&v)[NV], T *lds) {
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) lds[q * nw + wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            T acc = lds[q * nw];
            for (int w = 1; w < nw; w++) acc = acc + lds[q * nw + w];
            v[q] = acc;
        }
    }
}

// Reduce partials laid out [q * nparts + p] over p, for q < nv <= MAX_RED,
// with one block in one pass; deterministic.  Result in out[q] (shared
// memory), valid for all threads after return.
constexpr int MAX_RED = 68;   // >= SMAX + 4 (idrs.hpp)
constexpr int FIN_THREADS = 1024;   // block size of the one-block reduction kernels
template <class T>
__device__ inline void reduce_partials(const T *__restrict__ partials, int nparts, int nv, T *out) {
    __shared__ T lds[MAX_RED * (FIN_THREADS / WAVE)];
    const int nw = blockDim.x / WAVE, lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    for (int q = 0; q < nv; q++) {
        // four independent partial sums per thread keep several loads in flight
        T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
        const T *pq = partials + (long)q * nparts;
        int p = threadIdx.x;
        const int step = blockDim.x;
        for (; p + 3 * step < nparts; p += 4 * step) {
            a0 = a0 + pq[p];
            a1 = a1 + pq[p + step];
            a2 = a2 + pq[p + 2 * step];
            a3 = a3 + pq[p + 3 * step];
        }
        for (; p < nparts; p += step) a0 = a0 + pq[p];
        T v = wave_sum((a0 + a1) + (a2 + a3));
        if (lane == 0) lds[q * (FIN_THREADS / WAVE) + wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        const int q = threadIdx.x;
        T t = lds[q * (FIN_THREADS / WAVE)];
        for (int w = 1; w < nw; w++) t = t + lds[q * (FIN_THREADS / WAVE) + w];
        out[q] = t;
    }
    __syncthreads();
}

// Non-temporal loads for streams a kernel reads once and no later kernel
// finds in the caches (matrix values; the shadow space P and the older G
// and U columns, 8N bytes each, re-read only after other streams have
// passed): they should not push reused lines (the gathered x entries, the
// column just written) out of L2.  Non-temporal P loads made the 200^3
// P^T y product 4 % faster.
template <class V>
__device__ inline V ldnt(const V *p) { return __builtin_nontemporal_load(p); }
template <class R>
__device__ inline cplx<R> ldnt(const cplx<R> *p) {
    const R *q = reinterpret_cast<const R *>(p);
    return {__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1)};
}
// non-temporal stores, for outputs no later kernel reads soon
template <class V>
__device__ inline void stnt(V *p, V v) { __builtin_nontemporal_store(v, p); }
template <class R>
__device__ inline void stnt(cplx<R> *p, cplx<R> v) {
    R *q = reinterpret_cast<R *>(p);
    __builtin_nontemporal_store(v.re, q);
    __builtin_nontemporal_store(v.im, q + 1);
}
typedef double dbl2_nt __attribute__((ext_vector_type(2)));

// XCD-aware workgroup index: gfx950 dispatches consecutive workgroups round
// robin over its 8 XCDs, each with a private L2.  Remapping so that each XCD
// receives a contiguous range of logical blocks keeps neighbouring rows (and
// their shared gather targets) inside one L2.
constexpr int NUM_XCD = 8;
__device__ inline int xcd_block(int b, int nb) {
    if (nb % NUM_XCD != 0) return b;
    return (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
}

// splitmix64 (S. Vigna, public-domain generator) at a counter; top 53 bits
// to [0,1).  Default shadow-space stream.
__host__ __device__ inline double uniform01(uint64_t seed, uint64_t index) {
    uint64_t z = index + seed * 0xD1B54A32D192ED03ULL;
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace idrs
