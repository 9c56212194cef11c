// Mixed-precision IDR(s) (FP32 inner Krylov solve, FP64 defect correction).
//
// Each outer step: r = b - A x in FP64 (the true residual, A applied with
// FP64 values), stop if ||r|| <= tol ||b||; else solve A d = r / ||r|| by
// IDR(s) whose SpMVs read the FP32 copy of the (scaled) values -- with FP32
// vectors and scalars as well (fp32_vectors), or in FP64 otherwise -- to a
// relative tolerance inner_tol (no tighter than the remaining FP64 gap
// needs), and update x += ||r|| d in FP64.  The inner solve is the same
// device IDR(s) as the FP64 one (idrs.inc:1-311), instantiated for float /
// complex<float>.  The FP32 operator differs from A by O(1e-7) relative, so
// each outer step contracts the error by about max(inner_tol, cond * 1e-7).
#include "mixed.hpp"

#include <cmath>
#include <stdexcept>

namespace idrs {
namespace {

template <class T> struct Lo;
template <> struct Lo<double> { using type = float; };
template <> struct Lo<c128> { using type = c64; };

__device__ inline float lo(double v) { return (float)v; }
__device__ inline c64 lo(c128 v) { return c64((float)v.re, (float)v.im); }
__device__ inline double hi(float v) { return (double)v; }
__device__ inline c128 hi(c64 v) { return c128((double)v.re, (double)v.im); }

// r32 = r64 * c  (c = 1/||r||)
template <class T64, class T32>
__global__ void k_to_lo(long n, const T64 *__restrict__ r, double c, T32 *__restrict__ r32) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        if constexpr (Tr<T64>::is_complex) r32[q] = lo(T64(r[q].re * c, r[q].im * c));
        else r32[q] = lo(r[q] * c);
    }
}

// x = x + c * d32
template <class T64, class T32>
__global__ void k_add_hi(long n, T64 *__restrict__ x, double c, const T32 *__restrict__ d32) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const T64 d = hi(d32[q]);
        if constexpr (Tr<T64>::is_complex) x[q] = T64(fma(c, d.re, x[q].re), fma(c, d.im, x[q].im));
        else x[q] = fma(c, d, x[q]);
    }
}

// r = b - t, partials |r|^2
template <class T>
__global__ __launch_bounds__(256) void k_res(long n, const T *__restrict__ b, const T *__restrict__ t,
                                             T *__restrict__ r, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const T rq = t ? b[q] - t[q] : b[q];
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_sum(const T *partials, int nparts, T *out) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) out[0] = red[0];
}

// a device buffer from the solvers' caching pool (every call runs on the
// library's one stream, so a buffer handed back is reused in stream order)
struct Dev {
    void *p = nullptr;
    size_t n = 0;
    explicit Dev(size_t bytes) : n(bytes ? bytes : 1) { p = pool_alloc(n); }
    ~Dev() { pool_free(p, n); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    template <class V> V *as() const { return (V *)p; }
};

inline double real_part(double v) { return v; }
inline double real_part(c128 v) { return v.re; }

}  // namespace

// d64 = c * r64 (the FP64-vector inner solve's right-hand side)
template <class T>
__global__ void k_scale(long n, const T *__restrict__ r, double c, T *__restrict__ d) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        if constexpr (Tr<T>::is_complex) d[q] = T(r[q].re * c, r[q].im * c);
        else d[q] = r[q] * c;
    }
}

// x = x + c * d (FP64 correction)
template <class T>
__global__ void k_add(long n, T *__restrict__ x, double c, const T *__restrict__ d) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        if constexpr (Tr<T>::is_complex) x[q] = T(fma(c, d[q].re, x[q].re), fma(c, d[q].im, x[q].im));
        else x[q] = fma(c, d[q], x[q]);
    }
}

// the operator reads its FP32 values while in scope
struct Values32 {
    Operator &op;
    explicit Values32(Operator &o) : op(o) { op.values32 = true; }
    ~Values32() { op.values32 = false; }
};

template <class T64>
void mixed_solve(const Operator &A, long n, int s, const T64 *b, T64 *x, const SolveOpts<T64> &o, double inner_tol,
                 int max_outer, bool fp32_vectors, MixedStats &st, hipStream_t stream) {
    using T32 = typename Lo<T64>::type;
    if (n < 1) throw std::invalid_argument("empty system");
    if (A.K->nrows != n || A.K->ncols != n) throw std::invalid_argument("matrix is not n x n");
    if (A.general()) throw std::invalid_argument("mixed precision takes standard operators only");
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    if (!(inner_tol > 0 && inner_tol < 1)) throw std::invalid_argument("inner tolerance must be in (0, 1)");
    if (max_outer < 1) throw std::invalid_argument("at least one outer step");
    if (o.has_m1 || o.U0 || o.resvec || o.omega) throw std::invalid_argument("M1, U0, omega and resvec are not offered");
    if (A.values_complex()) throw std::invalid_argument("mixed precision needs real matrix values");
    if (A.seed != std::complex<double>(0.0, 0.0)) throw std::invalid_argument("mixed precision without a seed shift");
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * n, 1000);
    op_make_lowprec(const_cast<Operator &>(A));

    const size_t v64 = sizeof(T64) * (size_t)n, v32 = sizeof(T32) * (size_t)n;
    const int prow = o.variant == 1 ? s : 1;
    if (!fp32_vectors && !std::is_same<T64, double>::value)
        throw std::invalid_argument("FP32 values with FP64 vectors: real systems only");
    Dev br(v64), bt(v64), bw1(v64), bw2(v64), br32(v32), bd32(v32), bpart(sizeof(T64) * MAX_PARTS), bsum(sizeof(T64));
    std::unique_ptr<Dev> brh, bdh;
    if (!fp32_vectors) {
        brh.reset(new Dev(v64));
        bdh.reset(new Dev(v64));
    }
    std::unique_ptr<Dev> bP32;
    T64 *r = br.as<T64>(), *t = bt.as<T64>(), *w1 = bw1.as<T64>(), *w2 = bw2.as<T64>(), *part = bpart.as<T64>();
    T32 *r32 = br32.as<T32>(), *d32 = bd32.as<T32>();
    const int gv = vec_grid(n);
    EpiArgs<T64> none{};
    if (o.PT_raw) {
        bP32.reset(new Dev(v32 * prow));
        hipLaunchKernelGGL((k_to_lo<T64, T32>), dim3(vec_grid(n * prow)), dim3(256), 0, stream, n * (long)prow,
                           o.PT_raw, 1.0, bP32->as<T32>());
        LAUNCH_CHECK();
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    auto norm_of = [&](const T64 *bv, const T64 *tv, T64 *rv) {
        hipLaunchKernelGGL((k_res<T64>), dim3(gv), dim3(256), 0, stream, n, bv, tv, rv, part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_sum<T64>), dim3(1), dim3(FIN_THREADS), 0, stream, (const T64 *)part, gv, bsum.as<T64>());
        LAUNCH_CHECK();
        T64 h;
        HIP_TRY(hipMemcpyAsync(&h, bsum.p, sizeof(T64), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return std::sqrt(real_part(h));
    };
    const double normb = norm_of(b, nullptr, r);
    const double tolb = o.tol * normb;
    if (o.x0) {
        HIP_TRY(hipMemcpyAsync(x, o.x0, v64, hipMemcpyDeviceToDevice, stream));
        op_apply<T64>(A, x, t, w1, w2, EPI_NONE, none, nullptr, stream);
    } else {
        HIP_TRY(hipMemsetAsync(x, 0, v64, stream));
    }
    double normr = norm_of(b, o.x0 ? t : nullptr, r);
    int info = -1, its = 0, outer = 0;
    while (true) {
        if (normr <= tolb) {
            info = 0;
            break;
        }
        if (outer == max_outer || its >= maxit) {
            info = 1;
            break;
        }
        // no tighter than the FP64 target needs (half of it, for the FP32 error)
        const double itol = std::max(inner_tol, 0.5 * tolb / normr);
        SolveStats is;
        if (fp32_vectors) {
            hipLaunchKernelGGL((k_to_lo<T64, T32>), dim3(gv), dim3(256), 0, stream, n, (const T64 *)r, 1.0 / normr, r32);
            LAUNCH_CHECK();
            SolveOpts<T32> io;
            io.tol = itol;
            io.has_maxit = true;
            io.maxit = maxit - its;
            io.variant = o.variant;
            io.PT_raw = bP32 ? bP32->as<T32>() : nullptr;
            io.p_seed = o.p_seed;
            idrs_solve<T32>(A, n, s, r32, d32, io, is, stream);
            hipLaunchKernelGGL((k_add_hi<T64, T32>), dim3(gv), dim3(256), 0, stream, n, x, normr, (const T32 *)d32);
            LAUNCH_CHECK();
        } else {
            T64 *rh = brh->as<T64>(), *dh = bdh->as<T64>();
            hipLaunchKernelGGL((k_scale<T64>), dim3(gv), dim3(256), 0, stream, n, (const T64 *)r, 1.0 / normr, rh);
            LAUNCH_CHECK();
            SolveOpts<T64> io;
            io.tol = itol;
            io.has_maxit = true;
            io.maxit = maxit - its;
            io.variant = o.variant;
            io.PT_raw = o.PT_raw;
            io.p_seed = o.p_seed;
            {
                Values32 v32(const_cast<Operator &>(A));
                idrs_solve<T64>(A, n, s, rh, dh, io, is, stream);
            }
            hipLaunchKernelGGL((k_add<T64>), dim3(gv), dim3(256), 0, stream, n, x, normr, (const T64 *)dh);
            LAUNCH_CHECK();
        }
        its += is.iterations;
        outer++;
        op_apply<T64>(A, x, t, w1, w2, EPI_NONE, none, nullptr, stream);
        const double prev = normr;
        normr = norm_of(b, t, r);
        if (is.flag == 3 || !(normr < prev)) {     // inner breakdown, or no progress
            info = normr <= tolb ? 0 : 3;
            break;
        }
    }
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.flag = info;
    st.iterations = its;
    st.outer = outer;
    st.relres = normr / normb;
    st.solve_ms = ms;
}

// r = b - (t - sigma x), partials of |r|^2 (the shifted system's FP64 residual)
__global__ __launch_bounds__(256) void k_ms_res(long n, const c128 *__restrict__ b, const c128 *__restrict__ t,
                                                const c128 *__restrict__ x, c128 sigma, c128 *__restrict__ r,
                                                c128 *partials) {
    __shared__ c128 lds[4];
    c128 e[1] = {c128(0)};
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const c128 rq = b[q] - (t[q] - sigma * x[q]);
        r[q] = rq;
        e[0] = e[0] + c128(Tr<c128>::abs2(rq));
    }
    block_sum<c128, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

void ms_mixed_solve(const Operator &A, long n, int ns, const c128 *sigma, int s, const c128 *b, c128 *X,
                    const KrylovOpts<c128> &o, int method, double inner_tol, int max_outer, double *relres,
                    MsMixedStats &st, hipStream_t stream) {
    if (n < 1 || ns < 1) throw std::invalid_argument("empty system");
    if (A.K->nrows != n || A.K->ncols != n) throw std::invalid_argument("matrix is not n x n");
    if (A.general()) throw std::invalid_argument("mixed precision takes standard operators only");
    if (A.values_complex()) throw std::invalid_argument("mixed precision needs real matrix values");
    if (A.seed != std::complex<double>(0.0, 0.0)) throw std::invalid_argument("mixed precision without a seed shift");
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    if (!(inner_tol > 0 && inner_tol < 1)) throw std::invalid_argument("inner tolerance must be in (0, 1)");
    if (max_outer < 0) throw std::invalid_argument("max_outer must be >= 0");
    if (method != 0 && method != 1) throw std::invalid_argument("method: 0 MSQMRIDR, 1 MSIDRS");
    if (o.has_m1 || o.x0 || o.resvec || o.omega || o.want_colfac)
        throw std::invalid_argument("M1, x0, omega, resvec and colfac are not offered");
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * n, 1000);
    op_make_lowprec(const_cast<Operator &>(A));
    const size_t v64 = sizeof(c128) * (size_t)n, v32 = sizeof(c64) * (size_t)n;
    const int gv = vec_grid(n);
    Dev br(v64), bt(v64 * ns), bw1(v64 * ns), bw2(v64 * ns), bb32(v32), bX32(v32 * ns), bd32(v32),
        bpart(sizeof(c128) * MAX_PARTS), bsum(sizeof(c128));
    c128 *r = br.as<c128>(), *part = bpart.as<c128>();
    c64 *b32 = bb32.as<c64>(), *X32 = bX32.as<c64>(), *d32 = bd32.as<c64>();
    auto reduce = [&]() {
        hipLaunchKernelGGL((k_sum<c128>), dim3(1), dim3(FIN_THREADS), 0, stream, (const c128 *)part, gv, bsum.as<c128>());
        LAUNCH_CHECK();
        c128 h;
        HIP_TRY(hipMemcpyAsync(&h, bsum.p, sizeof(c128), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return std::sqrt(h.re);
    };
    // the complex64 solver's options: one solve of all shifts, or of one
    auto lo_opts = [&](double tol, int cap) {
        KrylovOpts<c64> io;
        io.tol = tol;
        io.has_maxit = true;
        io.maxit = cap;
        io.variant = o.variant;
        io.has_inner_s = o.has_inner_s;
        io.inner_s = o.inner_s;
        io.has_inner_tol = o.has_inner_tol;
        io.inner_tol = o.inner_tol;
        io.has_inner_it = o.has_inner_it;
        io.inner_it = o.inner_it;
        io.p_seed = o.p_seed;
        return io;
    };
    auto run_lo = [&](int nsh, const c64 *sig32, const c64 *rhs, c64 *out, double tol, int cap) {
        KrylovStats is;
        const KrylovOpts<c64> io = lo_opts(tol, cap);
        if (method == 0) msqmridr_solve<c64>(A, n, nsh, sig32, s, rhs, out, io, is, stream);
        else msidrs_solve<c64>(A, n, nsh, sig32, s, rhs, out, io, is, stream);
        return is;
    };
    std::vector<c64> sig32(ns);
    for (int k = 0; k < ns; k++) sig32[k] = c64((float)sigma[k].re, (float)sigma[k].im);

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    // ||b|| in FP64, b/||b|| in complex64
    hipLaunchKernelGGL((k_res<c128>), dim3(gv), dim3(256), 0, stream, n, b, (const c128 *)nullptr, r, part);
    LAUNCH_CHECK();
    const double normb = reduce();
    if (normb == 0.0) throw std::invalid_argument("zero right-hand side");
    const double tolb = o.tol * normb;
    hipLaunchKernelGGL((k_to_lo<c128, c64>), dim3(gv), dim3(256), 0, stream, n, b, 1.0 / normb, b32);
    LAUNCH_CHECK();
    // every shift at once in complex64
    KrylovStats s1 = run_lo(ns, sig32.data(), b32, X32, std::max(inner_tol, o.tol), maxit);
    int its = s1.iterations;
    HIP_TRY(hipMemsetAsync(X, 0, v64 * ns, stream));
    for (int k = 0; k < ns; k++) {
        hipLaunchKernelGGL((k_add_hi<c128, c64>), dim3(gv), dim3(256), 0, stream, n, X + (long)n * k, normb,
                           (const c64 *)(X32 + (long)n * k));
        LAUNCH_CHECK();
    }
    // FP64 residuals of every shift (block products of A on X)
    c128 *T = bt.as<c128>(), *W1 = bw1.as<c128>(), *W2 = bw2.as<c128>();
    std::vector<double> nr(ns);
    auto residual = [&](int k) {
        hipLaunchKernelGGL(k_ms_res, dim3(gv), dim3(256), 0, stream, n, b, (const c128 *)(T + (long)n * k),
                           (const c128 *)(X + (long)n * k), sigma[k], r, part);
        LAUNCH_CHECK();
        return reduce();
    };
    op_apply_cols<c128>(A, ns, n, X, T, W1, W2, stream);
    for (int k = 0; k < ns; k++) {
        nr[k] = residual(k);
        if (!(nr[k] <= normb)) {   // the complex64 solve broke down or diverged on this shift: restart it from zero
            HIP_TRY(hipMemsetAsync(X + (long)n * k, 0, v64, stream));
            nr[k] = normb;
        }
    }
    int outer = 0, corr = 0;
    bool stuck = false;
    while (!stuck) {
        bool all = true;
        for (int k = 0; k < ns; k++) all = all && nr[k] <= tolb;
        if (all || outer >= max_outer || its >= maxit) break;
        outer++;
        stuck = true;
        for (int k = 0; k < ns; k++) {
            if (nr[k] <= tolb || its >= maxit) continue;
            // correction of shift k in complex64: (A - sigma_k I) d = r_k / ||r_k||
            op_apply<c128>(A, X + (long)n * k, T + (long)n * k, W1, W2, EPI_NONE, EpiArgs<c128>{}, nullptr, stream);
            const double before = residual(k);
            hipLaunchKernelGGL((k_to_lo<c128, c64>), dim3(gv), dim3(256), 0, stream, n, (const c128 *)r, 1.0 / before,
                               b32);
            LAUNCH_CHECK();
            const double itol = std::max(inner_tol, 0.5 * tolb / before);
            KrylovStats sc = run_lo(1, &sig32[k], b32, d32, itol, maxit - its);
            its += sc.iterations;
            corr++;
            hipLaunchKernelGGL((k_add_hi<c128, c64>), dim3(gv), dim3(256), 0, stream, n, X + (long)n * k, before,
                               (const c64 *)d32);
            LAUNCH_CHECK();
            op_apply<c128>(A, X + (long)n * k, T + (long)n * k, W1, W2, EPI_NONE, EpiArgs<c128>{}, nullptr, stream);
            const double after = residual(k);
            if (after < before) {
                stuck = false;
                nr[k] = after;
            } else {                          // the correction diverged: take it back
                hipLaunchKernelGGL((k_add_hi<c128, c64>), dim3(gv), dim3(256), 0, stream, n, X + (long)n * k, -before,
                                   (const c64 *)d32);
                LAUNCH_CHECK();
                op_apply<c128>(A, X + (long)n * k, T + (long)n * k, W1, W2, EPI_NONE, EpiArgs<c128>{}, nullptr,
                               stream);
                nr[k] = residual(k);
            }
        }
    }
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    bool all = true;
    for (int k = 0; k < ns; k++) {
        relres[k] = nr[k] / normb;
        all = all && nr[k] <= tolb;
    }
    st.flag = all ? 0 : (stuck ? 3 : 1);
    st.iterations = its;
    st.outer = outer;
    st.corrections = corr;
    st.solve_ms = ms;
}

template void mixed_solve<double>(const Operator &, long, int, const double *, double *, const SolveOpts<double> &,
                                  double, int, bool, MixedStats &, hipStream_t);
template void mixed_solve<c128>(const Operator &, long, int, const c128 *, c128 *, const SolveOpts<c128> &, double,
                                int, bool, MixedStats &, hipStream_t);

}  // namespace idrs
