// QMRIDR, MSIDRS and MSQMRIDR on one GPU.
//
// Restates (paths relative to the reference's IDRS/):
//   QMRIDR    MODULES/IDRS/qmridr.inc:1-263   (RQMRIDR/CQMRIDR, idrs_module.f90:253-408)
//   MSIDRS    MODULES/IDRS/msidrs.inc:1-334   (RMSIDRS/CMSIDRS, idrs_module.f90:572-713)
//   MSQMRIDR  MODULES/IDRS/msqmridr.inc:1-293 (RMSQMRIDR/CMSQMRIDR, idrs_module.f90:410-570)
//   SOLVE, GROT  MODULES/IDRS/dense_la_module.f90:605-742
//
// Execution model: every vector operation and every reduction runs on the
// device; the O(s^2 * n_sigma) scalar recurrences (Hessenberg columns,
// per-shift LU pivots, Givens rotations, small GEPP solves) run on the host
// after each reduction, as the reference runs them on every image.  The
// per-shift updates of all shifts are fused into one kernel per step (shift
// loop inside the row loop: v and v0 are read once for all shifts).  The
// reference's column shifts of G and W (G(:,1:s-1) = G(:,2:s), ...) are
// circular windows here, so no vector is copied to age a column.
#include "krylov.hpp"
#include "pdot.hpp"

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace idrs {
namespace {

template <class T> struct HostT;
template <> struct HostT<double> { using type = double; };
template <> struct HostT<c128> { using type = std::complex<double>; };
// complex64 vectors (the multishift solvers' single-precision form) keep
// FP64 host scalars: the small recurrences, omegas and norms run in FP64
template <> struct HostT<c64> { using type = std::complex<double>; };

// host scalar -> device scalar of the vectors' type
template <class T, class H>
inline T to_devT(H v) {
    const std::complex<double> z(v);
    if constexpr (Tr<T>::is_complex) return T((typename Tr<T>::real)z.real(), (typename Tr<T>::real)z.imag());
    else return (T)z.real();
}
inline double to_host(double v) { return v; }
inline std::complex<double> to_host(c128 v) { return {v.re, v.im}; }
inline std::complex<double> to_host(c64 v) { return {(double)v.re, (double)v.im}; }
inline double conjh(double v) { return v; }
inline std::complex<double> conjh(std::complex<double> v) { return std::conj(v); }
inline double absh(double v) { return std::fabs(v); }
inline double absh(std::complex<double> v) { return std::abs(v); }
inline double reh(double v) { return v; }
inline double reh(std::complex<double> v) { return v.real(); }

#define KGRID(q, n) for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += (long)gridDim.x * blockDim.x)

constexpr int MAXT = 40;   // terms of one linear-combination launch
constexpr int MAXS = 64;   // shifts of one fused multishift update
// per-shift coefficients (s + 4 per shift) staged in LDS by the update
// kernels; larger sets (wide shadow spaces with many shifts) are read from
// device memory
constexpr int COEF_LDS = MAXS * (PDOT_MAX + 4);

__device__ inline double kdiv(double a, double b) { return a / b; }
__device__ inline c128 kdiv(c128 a, c128 b) { return a / b; }
__device__ inline c64 kdiv(c64 a, c64 b) { return a / b; }
template <class T>
__device__ inline T kfma(T a, T b, T c) { return Tr<T>::fma(a, b, c); }

template <class T>
struct Comb {
    int nt;
    const T *src[MAXT];
    T cof[MAXT];
};

// y = sum_t cof_t * src_t in term order (the first term multiplies, the rest
// accumulate), optional block partials of |y|^2.  y may alias a source.
template <class T>
__global__ __launch_bounds__(256) void k_comb(long n, T *y, Comb<T> c, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    KGRID(q, n) {
        T acc = c.cof[0] * c.src[0][q];
        for (int t = 1; t < c.nt; t++) acc = kfma(c.cof[t], c.src[t][q], acc);
        y[q] = acc;
        e[0] = e[0] + T(Tr<T>::abs2(acc));
    }
    if (partials) {
        block_sum<T, 1>(e, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
    }
}

// partials of sum conj(u_i) w_i for i < nu
template <class T>
struct Dots {
    int nu;
    const T *u[4];
    const T *w[4];
};
template <class T>
__global__ __launch_bounds__(256) void k_dots(long n, Dots<T> d, T *partials, int nparts) {
    __shared__ T lds[16];
    T e[4] = {T(0), T(0), T(0), T(0)};
    KGRID(q, n) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < d.nu) e[i] = kfma(Tr<T>::conj(d.u[i][q]), d.w[i][q], e[i]);
    }
    block_sum<T, 4>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < d.nu; i++) partials[(long)i * nparts + blockIdx.x] = e[i];
}

// row q of P (prow values, row-major) into pv: a whole, even-length row of
// doubles in 16-byte loads (every such row starts 16-byte aligned), else one
// value at a time (zero beyond prow)
template <class T, int NV>
__device__ inline void load_prow(const T *__restrict__ P, long q, int prow, T (&pv)[NV]) {
    if constexpr (std::is_same<T, double>::value && NV % 2 == 0) {
        if (prow == NV) {
#pragma unroll
            for (int i = 0; i < NV; i += 2) {
                const double2 v2 = *(const double2 *)(P + q * NV + i);
                pv[i] = v2.x;
                pv[i + 1] = v2.y;
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) pv[i] = i < prow ? P[q * prow + i] : T(0);
}

// partials of P^T w (no conjugation), prow rows
template <class T, int NV>
__global__ __launch_bounds__(256) void k_pdotk(long n, const T *__restrict__ w, const T *__restrict__ P, int prow,
                                               T *partials, int nparts) {
    __shared__ T lds[NV * 4];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    KGRID(q, n) {
        const T wq = w[q];
        T pv[NV];
        load_prow<T, NV>(P, q, prow, pv);
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) e[i] = kfma(pv[i], wq, e[i]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) partials[(long)i * nparts + blockIdx.x] = e[i];
}

// gn = v - om*t and partials of P^T gn  (msidrs.inc:165-168)
template <class T, int NV>
__global__ __launch_bounds__(256) void k_axpy_pdot(long n, T *gn, const T *__restrict__ v, const T *__restrict__ t,
                                                   T om, const T *__restrict__ P, int prow, T *partials, int nparts) {
    __shared__ T lds[NV * 4];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    KGRID(q, n) {
        const T g = kfma(-om, t[q], v[q]);
        gn[q] = g;
        T pv[NV];
        load_prow<T, NV>(P, q, prow, pv);
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) e[i] = kfma(pv[i], g, e[i]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) partials[(long)i * nparts + blockIdx.x] = e[i];
}

// gn = gn - v/om  (qmridr.inc:194)
template <class T>
__global__ void k_sub_div(long n, T *gn, const T *__restrict__ v, T om) {
    KGRID(q, n) gn[q] = gn[q] - kdiv(v[q], om);
}

// w = w / c for a real c
template <class T>
__global__ void k_div_real(long n, T *w, double c) {
    KGRID(q, n) {
        if constexpr (Tr<T>::is_complex) w[q] = T(w[q].re / c, w[q].im / c);
        else w[q] = w[q] / c;
    }
}

__device__ inline double real_part(double v) { return v; }
__device__ inline double real_part(c128 v) { return v.re; }
__device__ inline double real_part(c64 v) { return (double)v.re; }

// |w|^2 = the sum of the partials pin (reduced in every workgroup in one
// fixed order, so all hold the same value), then w = w / sqrt(|w|^2) with
// the square root taken on the device; workgroup 0 stores |w|^2 in nrm[0]
// and the root in nrm[1] (the host reads them with the other step scalars).
// The same pass writes the partials of P^T w for the normalised w (prow
// rows of P, no conjugation) to pout: the next step's P^T gn without
// another read of gn, with k_pdotk's grid, row order and block sums.
template <class T, int NV>
__global__ __launch_bounds__(256) void k_norm_div_pdot(long n, T *w, const T *__restrict__ pin, int npin, T *nrm,
                                                       const T *__restrict__ P, int prow, T *pout, int nparts) {
    __shared__ T red[1];
    __shared__ T lds[NV * 4];
    reduce_partials(pin, npin, 1, red);
    const double c = sqrt(real_part(red[0]));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        nrm[0] = red[0];
        nrm[1] = T(c);
    }
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    KGRID(q, n) {
        T g;
        if constexpr (Tr<T>::is_complex) g = T(w[q].re / c, w[q].im / c);
        else g = w[q] / c;
        w[q] = g;
        T pv[NV];
        load_prow<T, NV>(P, q, prow, pv);
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) e[i] = kfma(pv[i], g, e[i]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) pout[(long)i * nparts + blockIdx.x] = e[i];
}

// w = v / d (element-wise, d the M1 diagonal)
template <class T>
__global__ void k_div_vec(long n, const T *__restrict__ v, const T *__restrict__ d, T *__restrict__ w) {
    KGRID(q, n) w[q] = kdiv(v[q], d[q]);
}

// r = (b - t) + sigma*x, partials |r|^2 (msidrs.inc:318, msqmridr.inc:273-279)
template <class T>
__global__ __launch_bounds__(256) void k_shift_res(long n, const T *__restrict__ b, const T *__restrict__ t,
                                                   const T *__restrict__ x, T sigma, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    KGRID(q, n) {
        const T r = (b[q] - t[q]) + sigma * x[q];
        e[0] = e[0] + T(Tr<T>::abs2(r));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// Per-shift coefficients of the update kernels, passed by value in the
// launch when they fit (no copy to device memory before each update);
// larger sets come through device memory (the kernel's coef pointer).
template <class T>
struct CoefArg {
    static constexpr int CAP = 3072 / (int)sizeof(T);
    T c[CAP];
};

// MSIDRS dimension-reduction update of every shift (msidrs.inc:207-226):
//   t = v - om_s*v0*u(s+2-k) - sum_i W(:,i,s)*c_i;  W(:,k,s) = t/u(s+2);
//   x(:,s) = x(:,s) + phi_s*W(:,k,s)
// coef per shift: [om_s, u(s+2-k), c_1..c_s, u(s+2), phi_s]
// STAGED: the ns (s + 4) coefficients fit COEF_LDS and are staged in LDS;
// otherwise they are read from device memory (coef).
template <class T, bool STAGED>
__global__ __launch_bounds__(256) void k_msidrs_update(long n, int ns, int s, int k, const T *__restrict__ v,
                                                       const T *__restrict__ v0, T *W, T *X, const T *coef,
                                                       const CoefArg<T> ca) {
    __shared__ T sc[STAGED ? COEF_LDS : 1];
    const int stride = s + 4;
    if constexpr (STAGED) {
        for (int i = threadIdx.x; i < ns * stride; i += blockDim.x) sc[i] = coef ? coef[i] : ca.c[i];
        __syncthreads();
    }
    KGRID(q, n) {
        const T vq = v[q], v0q = v0[q];
        for (int ks = 0; ks < ns; ks++) {
            const T *c = (STAGED ? sc : coef) + ks * stride;
            T t = vq - (c[0] * v0q) * c[1];
            T *Wk = W + (long)n * s * ks;
            for (int i = 0; i < s; i++) t = kfma(-ldnt(&Wk[(long)n * i + q]), c[2 + i], t);
            const T wk = kdiv(t, c[s + 2]);
            stnt(&Wk[(long)n * k + q], wk);
            stnt(&X[(long)n * ks + q], kfma(c[s + 3], wk, ldnt(&X[(long)n * ks + q])));
        }
    }
}

// x(:,s) = x(:,s) + c_s*v0 for every shift (msidrs.inc:294)
template <class T>
__global__ void k_ms_xaxpy(long n, int ns, const T *__restrict__ v0, T *X, const T *coef, const CoefArg<T> ca) {
    __shared__ T sc[MAXS];
    for (int i = threadIdx.x; i < ns; i += blockDim.x) sc[i] = coef ? coef[i] : ca.c[i];
    __syncthreads();
    KGRID(q, n) {
        const T v0q = v0[q];
        for (int ks = 0; ks < ns; ks++) X[(long)n * ks + q] = kfma(sc[ks], v0q, X[(long)n * ks + q]);
    }
}

// QMR update of every active shift (qmridr.inc:238-243, msqmridr.inc:250-256):
//   wn = (vt_s - sum_j W(:,j,s) r_j)/r_{s+2};  W(:,oldest,s) = wn;  x_s += phi_s wn
// W columns are a circular window: logical column j (1..s+1) is physical
// (head + j - 1) mod (s+1).  coef per shift: [r_1..r_{s+1}, r_{s+2}, phi_s,
// active (real part != 0)].  vt_s is vt + ks*vt_stride.
template <class T, bool STAGED>
__global__ __launch_bounds__(256) void k_qmr_update(long n, int ns, int s, int head, const T *vt, long vt_stride,
                                                    T *W, T *X, const T *coef, const CoefArg<T> ca) {
    __shared__ T sc[STAGED ? COEF_LDS : 1];
    const int stride = s + 4, nw = s + 1;
    if constexpr (STAGED) {
        for (int i = threadIdx.x; i < ns * stride; i += blockDim.x) sc[i] = coef ? coef[i] : ca.c[i];
        __syncthreads();
    }
    KGRID(q, n) {
        for (int ks = 0; ks < ns; ks++) {
            const T *c = (STAGED ? sc : coef) + ks * stride;
            if (Tr<T>::absv(c[s + 3]) == 0.0) continue;
            T *Ws = W + (long)n * nw * ks;
            T m = T(0);
            for (int j = 0; j < nw; j++) {
                const int pj = (head + j) % nw;
                m = kfma(ldnt(&Ws[(long)n * pj + q]), c[j], m);
            }
            const T wn = kdiv(vt[vt_stride * ks + q] - m, c[s + 1]);
            stnt(&Ws[(long)n * head + q], wn);
            stnt(&X[(long)n * ks + q], kfma(c[s + 2], wn, ldnt(&X[(long)n * ks + q])));
        }
    }
}

// Copies nv device scalars into pinned host memory, then the call's
// sequence number (as k_red_host)
template <class T>
__global__ void k_fetch_host(const T *src, int nv, T *out, unsigned *seq_out, unsigned seq) {
    for (int i = threadIdx.x; i < nv; i += blockDim.x) out[i] = src[i];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// MGS step: beta = the sum of the previous dot's partials pin (reduced in
// every workgroup in one fixed order, so all hold the same value; workgroup
// 0 stores it in beta_out for the host); gn -= beta Gi, then the partials of
// Gn^H gn (or |gn|^2 without Gn) in pout (a buffer other than pin)
template <class T>
__global__ __launch_bounds__(256) void k_mgs_r(long n, T *gn, const T *__restrict__ Gi, const T *__restrict__ pin,
                                               int npin, T *beta_out, const T *__restrict__ Gn, T *partials) {
    __shared__ T lds[4];
    __shared__ T red[1];
    reduce_partials(pin, npin, 1, red);
    const T beta = red[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) *beta_out = beta;
    T e[1] = {T(0)};
    KGRID(q, n) {
        const T g = kfma(-beta, Gi[q], gn[q]);
        gn[q] = g;
        if (Gn) e[0] = kfma(Tr<T>::conj(Gn[q]), g, e[0]);
        else e[0] = e[0] + T(Tr<T>::abs2(g));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// Reduces the partials and writes the nv values straight into pinned host
// memory, then the call's sequence number: the host waits for that number
// instead of a copy and a stream synchronisation.  The values are made
// visible system-wide before the number (fence, then a release store).
template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_red_host(const T *partials, int nparts, int nv, T *out,
                                                          unsigned *seq_out, unsigned seq) {
    __shared__ T red[MAX_RED];
    reduce_partials(partials, nparts, nv, red);
    if (threadIdx.x < nv) out[threadIdx.x] = red[threadIdx.x];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// host-side workspace: reductions whose results land in pinned host memory
// ---------------------------------------------------------------------------
template <class T>
struct Work {
    using H = typename HostT<T>::type;
    static constexpr size_t NCOEF = (size_t)MAXS * (SMAX + 4);
    hipStream_t st;
    long n;
    int gv;
    T *part = nullptr, *part2 = nullptr, *dcoef = nullptr, *hcoef = nullptr, *dsc = nullptr;
    T *hres = nullptr, *hres_d = nullptr;      // coherent pinned result slots (host / device view)
    unsigned *hseq = nullptr, *hseq_d = nullptr;
    unsigned seq = 0;
    int cslot = 0;                             // hcoef half the next put_coef fills
    // device buffers (from the solvers' caching pool) and their sizes
    std::vector<std::pair<void *, size_t>> bufs;
    static constexpr unsigned MAPPED = hipHostMallocCoherent | hipHostMallocMapped;
    explicit Work(long n_, hipStream_t s) : st(s), n(n_), gv(vec_grid(n_)) {
        part = alloc(MAX_NV * (size_t)MAX_PARTS);
        part2 = alloc(MAX_PARTS);
        dcoef = alloc(NCOEF);
        dsc = alloc(MAX_RED);
        hres = (T *)pinned_take(sizeof(T) * 2 * MAX_RED, MAPPED);
        hseq = (unsigned *)pinned_take(64, MAPPED);
        *hseq = 0;
        HIP_TRY(hipHostGetDevicePointer((void **)&hres_d, hres, 0));
        HIP_TRY(hipHostGetDevicePointer((void **)&hseq_d, hseq, 0));
        hcoef = (T *)pinned_take(sizeof(T) * 2 * NCOEF, hipHostMallocDefault);
    }
    // back to the pools once the stream has finished with them
    ~Work() {
        (void)hipStreamSynchronize(st);
        for (auto &b : bufs) pool_free(b.first, b.second);
        pinned_give(hres, sizeof(T) * 2 * MAX_RED, MAPPED);
        pinned_give(hseq, 64, MAPPED);
        pinned_give(hcoef, sizeof(T) * 2 * NCOEF, hipHostMallocDefault);
    }
    // waits for k_red_host's sequence number; after a second of spinning
    // (long queued work) it waits on the stream instead
    void wait_seq(unsigned q) {
        const auto t0 = std::chrono::steady_clock::now();
        long spins = 0;
        while (__atomic_load_n(hseq, __ATOMIC_ACQUIRE) != q) {
            if ((++spins & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
                HIP_TRY(hipStreamSynchronize(st));
                if (__atomic_load_n(hseq, __ATOMIC_ACQUIRE) != q)
                    throw std::runtime_error("reduction result did not reach the host");
                break;
            }
        }
    }
    T *alloc(size_t count) {
        const size_t bytes = sizeof(T) * (count ? count : 1);
        void *p = pool_alloc(bytes);
        bufs.emplace_back(p, bytes);
        return (T *)p;
    }
    T *zeros(size_t count) {
        T *p = alloc(count);
        HIP_TRY(hipMemsetAsync(p, 0, sizeof(T) * count, st));
        return p;
    }
    // Posted transfers: each lands in its own region of the result slots
    // (off 0 or MAX_RED), and one wait on the last post's sequence number
    // covers all posts before it (stream order).
    unsigned post_fetch(const T *src, int nv, int off) {
        if (nv > MAX_RED) throw std::invalid_argument("too many scalars");
        const unsigned q = ++seq;
        hipLaunchKernelGGL((k_fetch_host<T>), dim3(1), dim3(64), 0, st, src, nv, hres_d + off, hseq_d, q);
        LAUNCH_CHECK();
        return q;
    }
    // nv sums of partials laid out [v][gv] to the result slots at off
    unsigned post_red(const T *partials, int nv, int off) {
        const unsigned q = ++seq;
        hipLaunchKernelGGL((k_red_host<T>), dim3(1), dim3(FIN_THREADS), 0, st, partials, gv, nv, hres_d + off, hseq_d,
                           q);
        LAUNCH_CHECK();
        return q;
    }
    std::vector<H> read(int nv, int off) const {
        std::vector<H> out(nv);
        for (int i = 0; i < nv; i++) out[i] = to_host(hres[off + i]);
        return out;
    }
    // reduce nv values of partials laid out [v][nparts] and bring them to the host
    std::vector<H> get(int nv, int nparts) {
        const unsigned q = ++seq;
        hipLaunchKernelGGL((k_red_host<T>), dim3(1), dim3(FIN_THREADS), 0, st, part, nparts, nv, hres_d, hseq_d, q);
        LAUNCH_CHECK();
        wait_seq(q);
        std::vector<H> out(nv);
        for (int i = 0; i < nv; i++) out[i] = to_host(hres[i]);
        return out;
    }
    // partials of u^H v into out (gv of them)
    void dot_parts(const T *u, const T *v, T *out) {
        Dots<T> d{};
        d.nu = 1;
        d.u[0] = u;
        d.w[0] = v;
        hipLaunchKernelGGL((k_dots<T>), dim3(gv), dim3(256), 0, st, n, d, out, gv);
        LAUNCH_CHECK();
    }
    // nv device scalars to the host
    std::vector<H> fetch(const T *src, int nv) {
        if (nv > MAX_RED) throw std::invalid_argument("too many scalars");
        const unsigned q = ++seq;
        hipLaunchKernelGGL((k_fetch_host<T>), dim3(1), dim3(64), 0, st, src, nv, hres_d, hseq_d, q);
        LAUNCH_CHECK();
        wait_seq(q);
        std::vector<H> out(nv);
        for (int i = 0; i < nv; i++) out[i] = to_host(hres[i]);
        return out;
    }
    // coefficients to the device; the two halves of hcoef alternate so a
    // half is rewritten only after the copy that read it has run (a get()
    // separates consecutive calls)
    void put_coef(const std::vector<H> &c) {
        if (c.size() > NCOEF) throw std::invalid_argument("too many coefficients");
        T *h = hcoef + (size_t)cslot * NCOEF;
        cslot ^= 1;
        for (size_t i = 0; i < c.size(); i++) h[i] = to_devT<T>(c[i]);
        HIP_TRY(hipMemcpyAsync(dcoef, h, sizeof(T) * c.size(), hipMemcpyHostToDevice, st));
    }
    // coefficients for an update kernel: in ca when they fit (nullptr
    // returned), else in device memory (its pointer returned)
    const T *coefs(const std::vector<H> &c, CoefArg<T> &ca) {
        if (c.size() <= (size_t)CoefArg<T>::CAP) {
            for (size_t i = 0; i < c.size(); i++) ca.c[i] = to_devT<T>(c[i]);
            return nullptr;
        }
        put_coef(c);
        return dcoef;
    }
    // y = sum c_i src_i; returns ||y|| when want_norm
    double comb(T *y, const std::vector<const T *> &src, const std::vector<H> &cof, bool want_norm) {
        size_t done = 0;
        double nrm = 0.0;
        // more terms than one launch holds: accumulate in chunks (y joins the next chunk)
        while (done < src.size()) {
            Comb<T> c{};
            int t = 0;
            if (done > 0) {
                c.src[t] = y;
                c.cof[t] = T(1.0);
                t++;
            }
            while (t < MAXT && done < src.size()) {
                c.src[t] = src[done];
                c.cof[t] = to_devT<T>(cof[done]);
                t++;
                done++;
            }
            c.nt = t;
            const bool last = done == src.size();
            hipLaunchKernelGGL((k_comb<T>), dim3(gv), dim3(256), 0, st, n, y, c, (last && want_norm) ? part : nullptr);
            LAUNCH_CHECK();
        }
        if (want_norm) nrm = std::sqrt(reh(get(1, gv)[0]));
        return nrm;
    }
    double norm(const T *w) {
        std::vector<const T *> s{w};
        T *tmp = scratch();
        return comb(tmp, s, {H(1.0)}, true);
    }
    T *scratch_ = nullptr;
    T *scratch() {
        if (!scratch_) scratch_ = alloc(n);
        return scratch_;
    }
    std::vector<H> dots(std::initializer_list<std::pair<const T *, const T *>> pairs) {
        Dots<T> d{};
        d.nu = 0;
        for (auto &p : pairs) {
            d.u[d.nu] = p.first;
            d.w[d.nu] = p.second;
            d.nu++;
        }
        hipLaunchKernelGGL((k_dots<T>), dim3(gv), dim3(256), 0, st, n, d, part, gv);
        LAUNCH_CHECK();
        return get(d.nu, gv);
    }
    void launch_pdotk(const T *w, const T *P, int prow) {
        if (prow > PDOT_MAX) return launch_pdot_wide<T>(n, w, P, prow, 0, part, 0, nullptr, st);
#define PD(NV) hipLaunchKernelGGL((k_pdotk<T, NV>), dim3(gv), dim3(256), 0, st, n, w, P, prow, part, gv)
        if (prow <= 1) PD(1);
        else if (prow <= 2) PD(2);
        else if (prow <= 4) PD(4);
        else if (prow <= 8) PD(8);
        else PD(16);
#undef PD
        LAUNCH_CHECK();
    }
    std::vector<H> pdot(const T *w, const T *P, int prow) {
        launch_pdotk(w, P, prow);
        return get(prow, gv);
    }
    std::vector<H> axpy_pdot(T *gn, const T *v, const T *t, H om, const T *P, int prow) {
#define AP(NV) hipLaunchKernelGGL((k_axpy_pdot<T, NV>), dim3(gv), dim3(256), 0, st, n, gn, v, t, to_devT<T>(om), P, prow, part, gv)
        if (prow > PDOT_MAX) {
            // gn here, P^T gn from the chunked kernel
            hipLaunchKernelGGL((k_axpy_pdot<T, 1>), dim3(gv), dim3(256), 0, st, n, gn, v, t, to_devT<T>(om), P, 0, part, gv);
            LAUNCH_CHECK();
            launch_pdot_wide<T>(n, gn, P, prow, 0, part, 0, nullptr, st);
            return get(prow, gv);
        }
        if (prow <= 1) AP(1);
        else if (prow <= 2) AP(2);
        else if (prow <= 4) AP(4);
        else if (prow <= 8) AP(8);
        else AP(16);
#undef AP
        LAUNCH_CHECK();
        return get(prow, gv);
    }
    void copy(T *dst, const T *src) {
        if (dst != src) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    }
};

// host SOLVE: Gaussian elimination with partial pivoting on the augmented
// matrix, back substitution (dense_la_module.f90:605-681)
template <class H>
std::vector<H> gepp(int n, const std::vector<H> &A /* column-major n x n */, const std::vector<H> &b) {
    std::vector<H> M((size_t)n * (n + 1));
    for (int c = 0; c < n; c++)
        for (int i = 0; i < n; i++) M[i + (size_t)n * c] = A[i + (size_t)n * c];
    for (int i = 0; i < n; i++) M[i + (size_t)n * n] = b[i];
    auto at = [&](int i, int c) -> H & { return M[i + (size_t)n * c]; };
    for (int j = 0; j < n - 1; j++) {
        int k = j;
        double best = absh(at(j, j));
        for (int i = j + 1; i < n; i++)
            if (absh(at(i, j)) > best) {
                best = absh(at(i, j));
                k = i;
            }
        if (k != j)
            for (int c = 0; c <= n; c++) std::swap(at(k, c), at(j, c));
        for (int i = j + 1; i < n; i++) {
            H p = at(i, j) / at(j, j);
            for (int c = j + 1; c <= n; c++) at(i, c) = at(i, c) - at(j, c) * p;
        }
    }
    std::vector<H> x(n);
    for (int i = n - 1; i >= 0; i--) {
        H p = at(i, n);
        for (int j = i + 1; j < n; j++) p = p - at(i, j) * at(j, n);
        at(i, n) = p / at(i, i);
    }
    for (int i = 0; i < n; i++) x[i] = at(i, n);
    return x;
}

// host GROT: Givens rotation in BLAS rotg style (dense_la_module.f90:683-742)
template <class H>
void grot(H a, H b, H &c, H &s, H &r) {
    if (absh(a) < DBL_EPSILON) {
        c = H(0.0);
        s = H(1.0);
        r = b;
    } else {
        const double rho = std::hypot(absh(a), absh(b));
        const H alpha = a / absh(a);
        c = H(absh(a) / rho);
        s = alpha * conjh(b) / rho;
        r = alpha * rho;
    }
}

// COMP_OM from (r,r), (t,t), (t,r) (idrs_module.f90:719-785)
template <class H>
H comp_om(H rr, H tt, H tr, double kappa) {
    if (kappa == 0.0) return tr / tt;
    const double rho = absh(tr) / std::sqrt(absh(rr * tt));
    H om = tr / tt;
    if (rho < kappa) om = om * kappa / rho;
    return om;
}

template <class T>
T *m1_diag(const KrylovOpts<T> &o, long n, Work<T> &w) {
    if (!o.has_m1 || !o.M1 || o.M1->which_precon != 1 || !o.M1->pdiag) return nullptr;
    if (o.M1->K->nrows != n) throw std::invalid_argument("M1 size does not match the system");
    T *d = w.alloc(n);
    if (o.M1->pdiag_kind == ValKind::Complex) {
        if constexpr (!Tr<T>::is_complex) throw std::invalid_argument("complex preconditioner for a real solve");
        else HIP_TRY(hipMemcpyAsync(d, o.M1->pdiag, sizeof(T) * n, hipMemcpyDeviceToDevice, w.st));
    } else {
        std::vector<double> r(n);
        HIP_TRY(hipMemcpyAsync(r.data(), o.M1->pdiag, sizeof(double) * n, hipMemcpyDeviceToHost, w.st));
        HIP_TRY(hipStreamSynchronize(w.st));
        std::vector<T> t(n);
        for (long i = 0; i < n; i++) t[i] = T(r[i]);
        HIP_TRY(hipMemcpyAsync(d, t.data(), sizeof(T) * n, hipMemcpyHostToDevice, w.st));
        HIP_TRY(hipStreamSynchronize(w.st));
    }
    return d;
}

template <class T>
void apply_m1(const Operator &A, const T *v, T *w, const T *m1, T *tmp, T *w1, T *w2, Work<T> &wk) {
    EpiArgs<T> none{};
    if (m1) {
        hipLaunchKernelGGL((k_div_vec<T>), dim3(wk.gv), dim3(256), 0, wk.st, wk.n, v, m1, tmp);
        LAUNCH_CHECK();
        op_apply<T>(A, tmp, w, w1, w2, EPI_NONE, none, nullptr, wk.st);
    } else {
        op_apply<T>(A, v, w, w1, w2, EPI_NONE, none, nullptr, wk.st);
    }
}

// True residual norms |b - (A x_k - sigma_k x_k)| of the shifted systems
// (A x = K'(x / M1) with M1), with A applied to the shifts' solutions as
// block products (one read of the matrix per 8 shifts) and one host read of
// the norms per block (msidrs.inc:312-320, msqmridr.inc:273-291).
template <class T>
void shifted_residuals(const Operator &A, long n, int ns, const std::vector<typename HostT<T>::type> &sigma,
                       const T *b, const T *X, const T *m1, Work<T> &w, std::vector<double> &normr) {
    const int nb = std::min(ns, SPMV_COLS);
    T *Tb = w.alloc((size_t)n * nb), *W1 = w.alloc((size_t)n * nb), *W2 = w.alloc((size_t)n * nb);
    T *Mb = m1 ? w.alloc((size_t)n * nb) : nullptr;
    for (int k0 = 0; k0 < ns; k0 += nb) {
        const int m = std::min(nb, ns - k0);
        const T *src = X + n * k0;
        if (m1) {
            for (int j = 0; j < m; j++) {
                hipLaunchKernelGGL((k_div_vec<T>), dim3(w.gv), dim3(256), 0, w.st, n, X + n * (k0 + j), m1, Mb + n * j);
                LAUNCH_CHECK();
            }
            src = Mb;
        }
        op_apply_cols<T>(A, m, n, src, Tb, W1, W2, w.st);
        for (int j = 0; j < m; j++) {
            hipLaunchKernelGGL((k_shift_res<T>), dim3(w.gv), dim3(256), 0, w.st, n, b, Tb + n * j, X + n * (k0 + j),
                               to_devT<T>(sigma[k0 + j]), w.part + (long)j * w.gv);
            LAUNCH_CHECK();
        }
        const std::vector<typename HostT<T>::type> r2 = w.get(m, w.gv);
        for (int j = 0; j < m; j++) normr[k0 + j] = std::sqrt(reh(r2[j]));
    }
}

inline int default_maxit(long n) { return (int)std::min<long>(2 * n, 1000); }

template <class T>
void check_common(const Operator &A, long n, int s, double tol) {
    if (s < 1 || s > SMAX) throw std::invalid_argument("s must be in 1..64");
    if (n < 1) throw std::invalid_argument("empty system");
    if (A.rows() != n || A.K->nrows != A.K->ncols) throw std::invalid_argument("matrix is not n x n");
    if (tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
}

}  // namespace

// ---------------------------------------------------------------------------
// MSIDRS (msidrs.inc)
// ---------------------------------------------------------------------------
template <class T>
void msidrs_solve(const Operator &A, long n, int ns, const T *sigma_d, int s, const T *b, T *X,
                  const KrylovOpts<T> &o, KrylovStats &out, hipStream_t stream) {
    using H = typename HostT<T>::type;
    check_common<T>(A, n, s, o.tol);
    if (ns < 1 || ns > MAXS) throw std::invalid_argument("number of shifts must be in 1..64");
    const int method = o.variant;
    if (method != 1 && method != 2) throw std::invalid_argument("variant must be 1 or 2");
    const int maxit = o.has_maxit ? o.maxit : default_maxit(n);
    if (o.resvec && maxit + 1 > o.resvec_len)
        throw std::invalid_argument("Length of vector with residual norms too small, should be maxit+1");
    const bool user_omega = o.omega && o.n_omega >= 1;
    const int prow = method == 1 ? s : 1;
    const double kappa = method == 1 ? (double)0.7f : 0.0;
    std::vector<H> sigma(ns);
    for (int k = 0; k < ns; k++) sigma[k] = to_host(sigma_d[k]);

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    Work<T> w(n, stream);
    T *G = w.zeros((size_t)n * s), *W = w.zeros((size_t)n * s * ns);
    // columns of G by pointer: storing gn as a column swaps buffers (gn is
    // rewritten before its next read) instead of copying
    std::vector<T *> Gc(s);
    for (int i = 0; i < s; i++) Gc[i] = G + (long)n * i;
    T *gn = w.alloc(n), *v = w.alloc(n), *v0 = w.zeros(n), *t = w.alloc(n);
    T *w1 = w.alloc(n), *w2 = w.alloc(n), *tmp = w.alloc(n), *P = w.alloc((size_t)n * prow);
    const T *m1 = m1_diag(o, n, w);
    HIP_TRY(hipMemsetAsync(X, 0, sizeof(T) * n * ns, stream));
    std::vector<double> normr(ns);
    auto rv = [&](int it) {
        if (o.resvec && it < o.resvec_len)
            for (int k = 0; k < ns; k++) o.resvec[it + (long)o.resvec_len * k] = normr[k];
    };
    const double normb = w.norm(b);                                       // msidrs.inc:70-75
    const double tolb = o.tol * normb;
    w.copy(gn, b);
    double normg = normb;
    for (auto &q : normr) q = normg;
    rv(0);
    int info = -1, iter = 0;
    std::vector<H> phi(ns, H(1.0)), oms(ns), l((size_t)(s + 2) * ns, H(0.0));
    if (normg <= tolb) {                                                  // :79-84
        info = 0;
    } else {
        make_p_device<T>(n, s, method, o.PT_raw, gn, P, o.p_seed, stream);   // :96
        std::vector<H> M((size_t)s * s, H(0.0)), f, alpha(s), gamma(s), h(s + 3), u(s + 2);
        auto MM = [&](int i, int j) -> H & { return M[(i - 1) + (size_t)s * (j - 1)]; };
        auto LS = [&](int i, int ks) -> H & { return l[(i - 1) + (size_t)(s + 2) * ks]; };
        H om = (double)sqrtf(2.0f);                                       // :103
        for (int k = 0; k < ns; k++) oms[k] = om / (1.0 - om * sigma[k]);
        int jj = 0;
        std::vector<H> coef((size_t)ns * (s + 4));
        while (info < 0) {
            f = w.pdot(gn, P, prow);                                      // :127
            f.resize(s, H(0.0));
            for (int k = 1; k <= s; k++) {
                iter = iter + 1;
                if (jj > 0) {                                             // :133-145
                    std::vector<const T *> src{gn};
                    std::vector<H> cof{H(1.0)};
                    for (int i = k; i <= s; i++) {
                        H g = f[i - 1];
                        for (int j = k; j <= i - 1; j++) g = g - MM(i, j) * gamma[j - 1];
                        g = g / MM(i, i);
                        gamma[i - 1] = g;
                        src.push_back(Gc[i - 1]);
                        cof.push_back(-g);
                    }
                    w.comb(v, src, cof, false);
                } else {
                    w.copy(v, gn);
                }
                for (int q = k; q <= s; q++) MM(q, k) = f[q - 1];         // :149
                if (absh(MM(k, k)) <= DBL_MIN) {                          // :151
                    info = 3;
                    break;
                }
                std::swap(Gc[k - 1], gn);                                 // G(:,k) = gn (:156)
                apply_m1<T>(A, v, t, m1, tmp, w1, w2, w);                 // :160-164
                f = w.axpy_pdot(gn, v, t, om, P, prow);                   // :165-168
                f.resize(s, H(0.0));
                std::vector<const T *> src{gn};
                std::vector<H> cof{H(1.0)};
                for (int i = 0; i < s; i++) alpha[i] = H(0.0);
                for (int i = 1; i <= k; i++) {                            // :170-178
                    H a = f[i - 1];
                    for (int j = 1; j <= i - 1; j++) a = a - MM(i, j) * alpha[j - 1];
                    a = a / MM(i, i);
                    alpha[i - 1] = a;
                    src.push_back(Gc[i - 1]);
                    cof.push_back(-a);
                    for (int q = k + 1; q <= s; q++) f[q - 1] = f[q - 1] - MM(q, i) * a;
                }
                normg = w.comb(gn, src, cof, true);                       // :179
                for (auto &q : h) q = H(0.0);                             // :183-188
                if (jj > 0)
                    for (int i = 1; i <= s - k + 1; i++) h[i - 1] = -gamma[k + i - 2];
                for (int i = 1; i <= k; i++) h[s + 2 - k + i - 1] = h[s + 2 - k + i - 1] - alpha[i - 1];
                h[s + 1] = h[s + 1] + 1.0;
                h[s + 2] = -1.0;
                for (auto &q : h) q = q / om;
                for (int ks = 0; ks < ns; ks++) {                         // :189-226
                    const H sg = sigma[ks];
                    for (int i = 0; i < s + 2; i++) u[i] = h[i];
                    if (jj > 0)
                        for (int i = 1; i <= s - k + 1; i++) u[i - 1] = u[i - 1] + sg * gamma[k + i - 2];
                    u[s + 1] = u[s + 1] - sg;
                    for (int i = 1; i <= s + 1; i++) LS(i, ks) = LS(i + 1, ks);
                    for (int j = 2; j <= s + 2; j++) u[j - 1] = u[j - 1] - LS(j - 1, ks) * u[j - 2];
                    LS(s + 2, ks) = -1.0 / (om * u[s + 1]);
                    H *c = &coef[(size_t)ks * (s + 4)];
                    c[0] = oms[ks];
                    c[1] = u[s + 2 - k - 1];
                    for (int i = 1; i <= s; i++) c[1 + i] = (i < k) ? u[s + 2 - k + i - 1] : u[i - k];
                    c[s + 2] = u[s + 1];
                    c[s + 3] = phi[ks];
                    phi[ks] = -LS(s + 2, ks) * phi[ks];
                    normr[ks] = normg * absh(phi[ks]);
                }
                CoefArg<T> ca{};
                const T *dc = w.coefs(coef, ca);
                if (ns * (s + 4) <= COEF_LDS)
                    hipLaunchKernelGGL((k_msidrs_update<T, true>), dim3(w.gv), dim3(256), 0, stream, n, ns, s, k - 1,
                                       v, v0, W, X, dc, ca);
                else
                    hipLaunchKernelGGL((k_msidrs_update<T, false>), dim3(w.gv), dim3(256), 0, stream, n, ns, s, k - 1,
                                       v, v0, W, X, dc, ca);
                LAUNCH_CHECK();
                rv(iter);
                double mx = 0.0;
                for (double q : normr) mx = std::max(mx, q);
                if (mx < tolb) {                                          // :230-236
                    info = 0;
                    break;
                } else if (iter == maxit) {
                    info = 1;
                    break;
                }
            }
            if (info >= 0) break;                                         // :242
            jj = jj + 1;
            iter = iter + 1;
            std::swap(v, gn);                                             // v = gn (:250)
            apply_m1<T>(A, v, t, m1, tmp, w1, w2, w);                     // :251-255
            if (user_omega) {                                             // :258-264
                int i = jj % o.n_omega;
                if (i == 0) i = o.n_omega;
                om = to_host(o.omega[i - 1]);
            } else {
                auto d = w.dots({{v, v}, {t, t}, {t, v}});
                om = comp_om<H>(d[0], d[1], d[2], kappa);
            }
            if (absh(om) <= DBL_EPSILON) {                                // :265-268
                info = 3;
                break;
            }
            normg = w.comb(gn, {v, t}, {H(1.0), -om}, true);              // :269-271
            std::swap(v0, v);                                             // v0 = v (:291)
            std::vector<H> xc(ns);
            for (int ks = 0; ks < ns; ks++) {                             // :280-299
                for (int i = 1; i <= s + 1; i++) LS(i, ks) = LS(i + 1, ks);
                LS(s + 2, ks) = -1.0 / (1.0 - om * sigma[ks]);
                oms[ks] = om / (1.0 - om * sigma[ks]);
                xc[ks] = phi[ks] * oms[ks];
                phi[ks] = -LS(s + 2, ks) * phi[ks];
                normr[ks] = normg * absh(phi[ks]);
            }
            CoefArg<T> ca{};
            const T *dc = w.coefs(xc, ca);
            hipLaunchKernelGGL((k_ms_xaxpy<T>), dim3(w.gv), dim3(256), 0, stream, n, ns, v0, X, dc, ca);
            LAUNCH_CHECK();
            rv(iter);
            double mx = 0.0;
            for (double q : normr) mx = std::max(mx, q);
            if (mx < tolb) {                                              // :301-307
                info = 0;
                break;
            } else if (iter == maxit) {
                info = 1;
                break;
            }
        }
    }
    shifted_residuals<T>(A, n, ns, sigma, b, X, m1, w, normr);            // :312-320
    if (m1 && !o.want_colfac)                                             // :323-327
        for (int ks = 0; ks < ns; ks++) {
            T *xk = X + (long)n * ks;
            hipLaunchKernelGGL((k_div_vec<T>), dim3(w.gv), dim3(256), 0, stream, n, xk, m1, tmp);
            LAUNCH_CHECK();
            w.copy(xk, tmp);
        }
    if (o.want_colfac && o.colfac)
        for (int ks = 0; ks < ns; ks++) o.colfac[ks] = to_devT<T>(phi[ks]);
    double mx = 0.0;
    for (double q : normr) mx = std::max(mx, q);
    if (info == 0 && mx >= tolb) info = 2;                                // :330
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out.flag = info;
    out.iterations = iter;
    out.relres = mx / normb;
    out.relres_vec.resize(ns);
    for (int ks = 0; ks < ns; ks++) out.relres_vec[ks] = normr[ks] / normb;
    out.solve_ms = ms;
}

// ---------------------------------------------------------------------------
// QMRIDR (qmridr.inc) and MSQMRIDR (msqmridr.inc) share one driver: the
// single-shift solver is the multishift one with sigma = 0 and its own
// preconditioner/inner-solver rules.
// ---------------------------------------------------------------------------
template <class T>
static void qmr_driver(const Operator &A, long n, int ns, const std::vector<typename HostT<T>::type> &sigma, int s,
                       const T *b, T *X, const KrylovOpts<T> &o, KrylovStats &out, hipStream_t stream, bool multishift) {
    using H = typename HostT<T>::type;
    check_common<T>(A, n, s, o.tol);
    if (ns < 1 || ns > MAXS) throw std::invalid_argument("number of shifts must be in 1..64");
    const int maxit = o.has_maxit ? o.maxit : default_maxit(n);
    int in_s = o.has_inner_s ? o.inner_s : (multishift ? 0 : 4);          // qmridr.inc:43, msqmridr.inc:45
    if (multishift ? in_s < 0 : in_s < 1) throw std::invalid_argument("Illegal value for parameter inner_s");
    const double in_tol = o.has_inner_tol ? o.inner_tol : (double)1.e-1f;
    if (in_tol < 0) throw std::invalid_argument("Illegal value for parameter inner_tolerance");
    const int in_it = o.has_inner_it ? o.inner_it : 0;
    if (multishift && in_it > 0 && in_s < 1) throw std::invalid_argument("inner iterations need inner_s >= 1");
    if (o.resvec && maxit + 1 > o.resvec_len)
        throw std::invalid_argument("Length of vector with residual norms too small, should be maxit+1");
    const bool user_omega = o.omega && o.n_omega >= 1;
    const double kappa = (double)0.7f;

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    Work<T> w(n, stream);
    const int nw = s + 1;
    T *G = w.zeros((size_t)n * s), *W = w.zeros((size_t)n * nw * ns);
    T *gn = w.alloc(n), *v = w.alloc(n), *t = w.alloc(n), *vt = w.alloc(n);
    T *w1 = w.alloc(n), *w2 = w.alloc(n), *tmp = w.alloc(n), *P = w.alloc((size_t)n * s);
    T *Vt = nullptr;
    if (multishift && in_it > 0) Vt = w.alloc((size_t)n * (ns + 1));
    const T *m1 = m1_diag(o, n, w);
    std::vector<double> normr(ns);
    auto rv = [&](int it) {
        if (o.resvec && it < o.resvec_len)
            for (int k = 0; k < ns; k++) o.resvec[it + (long)o.resvec_len * k] = normr[k];
    };
    EpiArgs<T> none{};
    // x = 0 or x0; gn = b - A x (qmridr.inc:65-84; msqmridr.inc:66-86)
    const double normb = w.norm(b);
    const double tolb = o.tol * normb;
    HIP_TRY(hipMemsetAsync(X, 0, sizeof(T) * n * ns, stream));
    if (!multishift && o.x0) {
        w.copy(X, o.x0);
        op_apply<T>(A, X, t, w1, w2, EPI_NONE, none, nullptr, stream);
        w.comb(gn, {b, t}, {H(1.0), H(-1.0)}, false);
    } else {
        w.copy(gn, b);
    }
    double normg = (!multishift && o.x0) ? w.norm(gn) : normb;
    for (auto &q : normr) q = normg;
    rv(0);
    int info = -1, iter = 0;
    if (normg <= tolb) {
        info = 0;
    } else {
        make_p_device<T>(n, s, 1, o.PT_raw, gn, P, o.p_seed, stream);     // :108
        hipLaunchKernelGGL((k_div_real<T>), dim3(w.gv), dim3(256), 0, stream, n, gn, normg);   // :112
        LAUNCH_CHECK();
        std::vector<H> M((size_t)s * s, H(0.0)), gamma(s), c(s + 2), h(s + 2), rs(s + 3), mu;
        bool mu_ready = false;
        std::vector<H> cs((size_t)(s + 2) * ns, H(0.0)), sn((size_t)(s + 2) * ns, H(0.0));
        std::vector<H> phn(ns, H(0.0)), phn1(ns, H(normg)), eta(ns, H(1.0));
        std::vector<char> converged(ns, 0);
        auto MQ = [&](int i, int j) -> H & { return M[(i - 1) + (size_t)s * (j - 1)]; };
        auto CS = [&](int i, int ks) -> H & { return cs[(i - 1) + (size_t)(s + 2) * ks]; };
        auto SN = [&](int i, int ks) -> H & { return sn[(i - 1) + (size_t)(s + 2) * ks]; };
        H om = multishift ? H((double)sqrtf(2.0f)) : H(1.0);              // qmridr.inc:116, msqmridr.inc:117
        int jj = 0, ghead = 0, whead = 0;
        // columns of G by pointer (see msidrs_solve): logical column i (1-based)
        std::vector<T *> Gc(s);
        for (int i = 0; i < s; i++) Gc[i] = G + (long)n * i;
        auto Gcol = [&](int i) { return Gc[(ghead + i - 1) % s]; };
        std::vector<H> coef((size_t)ns * (s + 4));
        std::vector<T> sig0;
        if (multishift && in_it > 0) {
            sig0.resize(ns + 1);
            for (int k = 0; k < ns; k++) sig0[k] = to_devT<T>(sigma[k]);
            sig0[ns] = T(0.0);
        }
        while (info < 0) {
            for (int k = 1; k <= s + 1; k++) {
                iter = iter + 1;
                for (auto &q : c) q = H(0.0);
                c[s] = H(1.0);
                // P^T gn (:139-140): after the first step it came back with
                // the previous step's norm
                if (!mu_ready) mu = w.pdot(gn, P, s);
                mu_ready = false;
                if (iter > s) {
                    gamma = gepp<H>(s, M, mu);                            // :144-150
                    std::vector<const T *> src{gn};
                    std::vector<H> cof{H(1.0)};
                    for (int i = 1; i <= s; i++) {
                        src.push_back(Gcol(i));
                        cof.push_back(-gamma[i - 1]);
                    }
                    w.comb(v, src, cof, false);
                    for (int i = 0; i < s; i++) c[i] = -gamma[i];
                } else {
                    w.copy(v, gn);
                }
                for (int j = 1; j <= s - 1; j++)                          // :155-158
                    for (int i = 1; i <= s; i++) MQ(i, j) = MQ(i, j + 1);
                for (int i = 1; i <= s; i++) MQ(i, s) = mu[i - 1];
                std::swap(Gc[ghead], gn);                                 // oldest slot becomes column s
                ghead = (ghead + 1) % s;
                const T *vtil = vt;
                if (in_it == 0) {                                         // :161-175
                    if (!multishift && m1) {
                        hipLaunchKernelGGL((k_div_vec<T>), dim3(w.gv), dim3(256), 0, stream, n, v, m1, vt);
                        LAUNCH_CHECK();
                    } else {
                        vtil = v;
                    }
                } else if (!multishift) {
                    SolveOpts<T> io;
                    io.has_m1 = o.has_m1;
                    io.M1 = o.M1;
                    io.tol = in_tol;
                    io.has_maxit = true;
                    io.maxit = in_it;
                    io.p_seed = 1;
                    SolveStats ist;
                    idrs_solve<T>(A, n, in_s, v, vt, io, ist, stream);
                } else {
                    KrylovOpts<T> io;
                    io.has_m1 = o.has_m1;
                    io.M1 = o.M1;
                    io.tol = in_tol;
                    io.has_maxit = true;
                    io.maxit = in_it;
                    io.p_seed = 1;
                    std::vector<T> cf(ns + 1);
                    io.colfac = cf.data();
                    io.want_colfac = true;
                    KrylovStats ist;
                    msidrs_solve<T>(A, n, ns + 1, sig0.data(), in_s, v, Vt, io, ist, stream);
                    for (int k2 = 0; k2 < ns; k2++) eta[k2] = to_host(cf[k2]) / to_host(cf[ns]);
                    vtil = Vt + (long)n * ns;
                }
                if (multishift) apply_m1<T>(A, vtil, gn, m1, tmp, w1, w2, w);   // :178-182
                else op_apply<T>(A, vtil, gn, w1, w2, EPI_NONE, none, nullptr, stream);
                if (k == s + 1) {                                         // :179-193
                    if (user_omega) {
                        int i = jj % o.n_omega;
                        if (i == 0) i = o.n_omega;
                        om = to_host(o.omega[i - 1]);
                    } else {
                        auto d = w.dots({{v, v}, {gn, gn}, {gn, v}});
                        om = comp_om<H>(d[0], d[1], d[2], kappa);
                    }
                    if (absh(om) <= DBL_EPSILON) {
                        info = 3;
                        break;
                    }
                    jj = jj + 1;
                }
                hipLaunchKernelGGL((k_sub_div<T>), dim3(w.gv), dim3(256), 0, stream, n, gn, v, to_devT<T>(om));  // :194
                LAUNCH_CHECK();
                for (int i = 0; i < s + 2; i++) h[i] = c[i] / om;         // :197
                // The step's scalars stay on the device until one host read:
                // the MGS betas db[i], i = i0..s (:198-205, modified
                // Gram-Schmidt; none in the last step), |gn|^2 in db[s+1],
                // normg = sqrt of it in db[s+2] (gn /= normg on the device,
                // :206-208).  The next step's P^T gn is posted behind them,
                // so both come back in one round trip.
                {
                    T *db = w.dsc;
                    const int i0 = k < s + 1 ? s - k + 1 : s + 1;
                    // each kernel reduces the partials of the dot before it
                    // (two partial buffers in turn)
                    T *pa = w.part, *pb = w.part2;
                    if (k < s + 1) {
                        w.dot_parts(Gcol(i0), gn, pa);
                        for (int i = i0; i <= s; i++) {
                            const T *Gn = i < s ? Gcol(i + 1) : nullptr;
                            hipLaunchKernelGGL((k_mgs_r<T>), dim3(w.gv), dim3(256), 0, stream, n, gn, Gcol(i), pa, w.gv,
                                               db + i, Gn, pb);
                            LAUNCH_CHECK();
                            std::swap(pa, pb);
                        }
                    } else {
                        w.dot_parts(gn, gn, pa);
                    }
                    // gn /= |gn| fused with the next step's P^T gn partials (in pb)
#define NDP(NV)                                                                                                     \
    hipLaunchKernelGGL((k_norm_div_pdot<T, NV>), dim3(w.gv), dim3(256), 0, stream, n, gn, pa, w.gv, db + s + 1, P, s, \
                       pb, w.gv)
                    if (s > PDOT_MAX) {
                        // gn /= |gn| here, then P^T gn from the chunked kernel
                        hipLaunchKernelGGL((k_norm_div_pdot<T, 1>), dim3(w.gv), dim3(256), 0, stream, n, gn, pa, w.gv,
                                           db + s + 1, P, 0, pb, w.gv);
                        LAUNCH_CHECK();
                        launch_pdot_wide<T>(n, gn, P, s, 0, pb, 0, nullptr, stream);
                    } else if (s <= 1) NDP(1);
                    else if (s <= 2) NDP(2);
                    else if (s <= 4) NDP(4);
                    else if (s <= 8) NDP(8);
                    else NDP(16);
#undef NDP
                    LAUNCH_CHECK();
                    w.post_fetch(db + i0, s + 3 - i0, 0);
                    w.wait_seq(w.post_red(pb, s, MAX_RED));
                    const std::vector<H> bv = w.read(s + 3 - i0, 0);
                    for (int i = i0; i <= s; i++) h[i] = h[i] + bv[i - i0];
                    normg = reh(bv[s + 2 - i0]);
                    mu = w.read(s, MAX_RED);
                    mu_ready = true;
                }
                h[s + 1] = normg;
                bool broke = false;
                for (int ks = 0; ks < ns; ks++) {                         // :211-243 / :217-259
                    H *cc = &coef[(size_t)ks * (s + 4)];
                    if (converged[ks]) {
                        cc[s + 3] = H(0.0);
                        continue;
                    }
                    for (auto &q : rs) q = H(0.0);
                    for (int i = 0; i < s + 2; i++) {
                        if (!multishift) rs[i + 1] = h[i];
                        else if (in_it == 0) rs[i + 1] = h[i] - sigma[ks] * c[i];
                        else rs[i + 1] = eta[ks] * h[i] - (eta[ks] - 1.0) * c[i];
                    }
                    const int low = std::max(1, s + 3 - iter);
                    for (int l = low; l <= s + 1; l++) {
                        H tau = rs[l - 1];
                        rs[l - 1] = CS(l, ks) * tau + SN(l, ks) * rs[l];
                        rs[l] = -conjh(SN(l, ks)) * tau + CS(l, ks) * rs[l];
                    }
                    {
                        H tau = rs[s + 1];
                        grot<H>(tau, rs[s + 2], CS(s + 2, ks), SN(s + 2, ks), rs[s + 1]);
                    }
                    phn[ks] = CS(s + 2, ks) * phn1[ks];
                    phn1[ks] = -conjh(SN(s + 2, ks)) * phn1[ks];
                    for (int i = 1; i <= s + 1; i++) {
                        CS(i, ks) = CS(i + 1, ks);
                        SN(i, ks) = SN(i + 1, ks);
                    }
                    if (absh(rs[s + 1]) < DBL_EPSILON) {
                        info = 3;
                        broke = true;
                        for (int k2 = ks; k2 < ns; k2++) coef[(size_t)k2 * (s + 4) + s + 3] = H(0.0);
                        break;
                    }
                    for (int i = 0; i < s + 1; i++) cc[i] = rs[i];
                    cc[s + 1] = rs[s + 1];
                    cc[s + 2] = phn[ks];
                    cc[s + 3] = H(1.0);
                    normr[ks] = absh(phn1[ks]) * (double)sqrtf((float)(jj + 1));
                    if (multishift) converged[ks] = normr[ks] < tolb;
                }
                CoefArg<T> ca{};
                const T *dc = w.coefs(coef, ca);
                const T *vsrc = (multishift && in_it > 0) ? Vt : vtil;
                const long vstride = (multishift && in_it > 0) ? n : 0;
                if (ns * (s + 4) <= COEF_LDS)
                    hipLaunchKernelGGL((k_qmr_update<T, true>), dim3(w.gv), dim3(256), 0, stream, n, ns, s, whead,
                                       vsrc, vstride, W, X, dc, ca);
                else
                    hipLaunchKernelGGL((k_qmr_update<T, false>), dim3(w.gv), dim3(256), 0, stream, n, ns, s, whead,
                                       vsrc, vstride, W, X, dc, ca);
                LAUNCH_CHECK();
                whead = (whead + 1) % nw;
                if (!multishift && broke) break;                          // qmridr.inc:234-236
                rv(iter);
                double mx = 0.0;
                for (double q : normr) mx = std::max(mx, q);
                if (mx < tolb) {
                    info = 0;
                    break;
                } else if (iter == maxit) {
                    info = 1;
                    break;
                }
            }
        }
    }
    // final true residuals (qmridr.inc:256-259; msqmridr.inc:273-291)
    if (multishift) {
        shifted_residuals<T>(A, n, ns, sigma, b, X, m1, w, normr);
    } else {
        op_apply<T>(A, X, t, w1, w2, EPI_NONE, none, nullptr, stream);
        hipLaunchKernelGGL((k_shift_res<T>), dim3(w.gv), dim3(256), 0, stream, n, b, t, X, T(0), w.part);
        LAUNCH_CHECK();
        normr[0] = std::sqrt(reh(w.get(1, w.gv)[0]));
    }
    if (multishift && m1)
        for (int ks = 0; ks < ns; ks++) {
            T *xk = X + (long)n * ks;
            hipLaunchKernelGGL((k_div_vec<T>), dim3(w.gv), dim3(256), 0, stream, n, xk, m1, tmp);
            LAUNCH_CHECK();
            w.copy(xk, tmp);
        }
    double mx = 0.0;
    for (double q : normr) mx = std::max(mx, q);
    if (info == 0 && mx >= tolb) info = 2;
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out.flag = info;
    out.iterations = iter;
    out.relres = mx / normb;
    out.relres_vec.resize(ns);
    for (int ks = 0; ks < ns; ks++) out.relres_vec[ks] = normr[ks] / normb;
    out.solve_ms = ms;
}

template <class T>
void qmridr_solve(const Operator &A, long n, int s, const T *b, T *x, const KrylovOpts<T> &o, KrylovStats &st,
                  hipStream_t stream) {
    std::vector<typename HostT<T>::type> sig(1, 0.0);
    qmr_driver<T>(A, n, 1, sig, s, b, x, o, st, stream, false);
}

template <class T>
void msqmridr_solve(const Operator &A, long n, int ns, const T *sigma, int s, const T *b, T *X,
                    const KrylovOpts<T> &o, KrylovStats &st, hipStream_t stream) {
    std::vector<typename HostT<T>::type> sig(ns);
    for (int k = 0; k < ns; k++) sig[k] = to_host(sigma[k]);
    qmr_driver<T>(A, n, ns, sig, s, b, X, o, st, stream, true);
}

#define KINST(T)                                                                                                \
    template void qmridr_solve<T>(const Operator &, long, int, const T *, T *, const KrylovOpts<T> &,           \
                                  KrylovStats &, hipStream_t);                                                  \
    template void msidrs_solve<T>(const Operator &, long, int, const T *, int, const T *, T *,                  \
                                  const KrylovOpts<T> &, KrylovStats &, hipStream_t);                           \
    template void msqmridr_solve<T>(const Operator &, long, int, const T *, int, const T *, T *,                \
                                    const KrylovOpts<T> &, KrylovStats &, hipStream_t);
KINST(double)
KINST(c128)
KINST(c64)
#undef KINST

}  // namespace idrs
