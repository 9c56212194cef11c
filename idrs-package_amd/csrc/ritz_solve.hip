// IDRS_RITZ (MODULES/IDRS/ritz_module.f90:225-556): IDR(s) with its
// Hessenberg matrix, omega suggestions from the Ritz values, and the recycle
// space U0 rebuilt on the device.
//
// The solve and the Krylov basis recurrence of compute_U0
// (ritz_module.f90:100-223) run on the GPU: one product per basis vector and
// one fused streaming kernel that forms the next vector from the product and
// the ring of the last dimU0 + 1 vectors, stores it, and adds it into the
// dimU0 columns of U0.  The small dense work (QR eigenvalues, shift-and-invert
// power method, omega ordering) runs on the host in ritz.cpp.  The global QR
// factorisation (ritz_module.f90:36-98) is modified Gram-Schmidt over device
// columns with deterministic two-level sums.
#include "idrs.hpp"
#include "ritz.hpp"

#include <algorithm>
#include <complex>
#include <limits>
#include <stdexcept>
#include <vector>

namespace idrs {

namespace {

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t bytes = 0;
    explicit DBuf(size_t cnt) : bytes(sizeof(T) * (cnt ? cnt : 1)) { p = (T *)pool_alloc(bytes); }
    ~DBuf() { pool_free(p, bytes); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

inline std::complex<double> to_std(double v) { return v; }
inline std::complex<double> to_std(c128 v) { return {v.re, v.im}; }
template <class T> inline T from_std(std::complex<double> v);
template <> inline c128 from_std<c128>(std::complex<double> v) { return c128(v.real(), v.imag()); }

constexpr int RB = 256;   // block size of the streaming kernels

#define RGRID(q, n) for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += (long)gridDim.x * blockDim.x)

// W(:, slot) = v = b; U0(:, j) = U0(:, j) + vec(1, j) v over U0 = 0
// (ritz_module.f90:168-178, i = 1)
template <class T>
__global__ __launch_bounds__(RB) void k_u0_first(long n, int d, const T *__restrict__ b, T *__restrict__ W,
                                                 T *__restrict__ U0, const T *__restrict__ vecrow) {
#pragma clang fp contract(off)
    RGRID(q, n) {
        const T v = b[q];
        W[q] = v;
        for (int j = 0; j < d; j++) U0[q + (long)n * j] = T(0) + vecrow[j] * v;
    }
}

// v = (A v_i - sum_{j=low..i} W(:, ind(j)) H(j, i)) / H(i+1, i), the sum in
// ascending j (ritz_module.f90:180-184); then, unless this was the last
// basis vector, W(:, ind(i+1)) = v and U0(:, j) += vec(i+1, j) v
// (ritz_module.f90:175-178).  hc holds H(low..i+1, i); slots the ring
// columns of low..i; out the ring column of i+1 (it may be slot[0]: each
// row is read before it is written).
template <class T>
__global__ __launch_bounds__(RB) void k_u0_step(long n, int d, int nlow, const T *__restrict__ t, T *W,
                                                const int *__restrict__ slots, const T *__restrict__ hc, int out,
                                                T *__restrict__ U0, const T *__restrict__ vecrow) {
#pragma clang fp contract(off)
    RGRID(q, n) {
        T acc = W[q + (long)n * slots[0]] * hc[0];
        for (int j = 1; j < nlow; j++) acc = acc + W[q + (long)n * slots[j]] * hc[j];
        const T v = (t[q] - acc) / hc[nlow];
        if (vecrow) {
            W[q + (long)n * out] = v;
            for (int j = 0; j < d; j++) U0[q + (long)n * j] = U0[q + (long)n * j] + vecrow[j] * v;
        }
    }
}

// partials[block] = sum conj(x) y over the block's rows
template <class T>
__global__ __launch_bounds__(RB) void k_dot_partials(long n, const T *__restrict__ x, const T *__restrict__ y,
                                                     T *partials) {
    __shared__ T lds[RB / WAVE];
    T e[1] = {T(0)};
    RGRID(q, n) e[0] = e[0] + Tr<T>::conj(x[q]) * y[q];
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_dot_fin(const T *partials, int nparts, T *out) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) out[0] = red[0];
}

// y = y - a x (a from the device), or x = x / nrm
template <class T>
__global__ __launch_bounds__(RB) void k_sub_scaled(long n, const T *a, const T *__restrict__ x, T *__restrict__ y) {
#pragma clang fp contract(off)
    const T s = a[0];
    RGRID(q, n) y[q] = y[q] - s * x[q];
}
template <class T>
__global__ __launch_bounds__(RB) void k_div_real(long n, double nrm, T *x) {
    RGRID(q, n) x[q] = x[q] / nrm;
}

}  // namespace

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
namespace {

template <class T>
struct Reducer {
    int g;
    DBuf<T> part, out;
    explicit Reducer(long n) : g(std::min(vec_grid(n), 4096)), part(g), out(1) {}
    // conj(x)^T y, on the host
    T dot(long n, const T *x, const T *y, hipStream_t st) {
        hipLaunchKernelGGL((k_dot_partials<T>), dim3(g), dim3(RB), 0, st, n, x, y, part.p);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_dot_fin<T>), dim3(1), dim3(FIN_THREADS), 0, st, part.p, g, out.p);
        LAUNCH_CHECK();
        T h;
        HIP_TRY(hipMemcpyAsync(&h, out.p, sizeof(T), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return h;
    }
    double norm(long n, const T *x, hipStream_t st) { return std::sqrt(std::real(to_std(dot(n, x, x, st)))); }
};

}  // namespace

// compute_U0 (ritz_module.f90:100-223) into U (n x d, device): the Krylov
// basis of b regenerated from H(1:nritz+1, 1:nritz) (host, leading
// dimension ldh), combined with vec (nritz x d, host, column-major); the
// columns normalised (the reference normalises when resnrm is requested,
// which IDRS_RITZ always does; the residual norms themselves are not
// returned by IDRS_RITZ).  M: v/M before each product when present.
template <class T>
static void compute_u0(const Operator &A, const Operator *M, long n, const T *b, const T *H, long ldh, int nritz,
                       const T *vec, int d, T *U, hipStream_t st) {
    const int g = vec_grid(n);
    DBuf<T> W((size_t)n * (d + 1)), t(n), w1(n), w2(n), tmp(M ? n : 1);
    // per step: the band of H's column and the next row of vec
    std::vector<int> slots_h((size_t)nritz * (d + 1));
    std::vector<T> hc_h((size_t)nritz * (d + 2)), vec_h((size_t)(nritz + 1) * d, T(0));
    std::vector<int> ind(nritz + 1);
    int k = 0;
    for (int i = 1; i <= nritz; i++) {
        k = k + 1;
        if (k > d + 1) k = 1;
        ind[i - 1] = k - 1;
    }
    for (int i = 1; i <= nritz; i++) {
        const int low = std::max(1, i - d);
        for (int j = low; j <= i; j++) {
            slots_h[(size_t)(i - 1) * (d + 1) + (j - low)] = ind[j - 1];
            hc_h[(size_t)(i - 1) * (d + 2) + (j - low)] = H[(j - 1) + ldh * (i - 1)];
        }
        hc_h[(size_t)(i - 1) * (d + 2) + (i - low + 1)] = H[i + ldh * (i - 1)];
        for (int j = 0; j < d; j++) vec_h[(size_t)(i - 1) * d + j] = vec[(i - 1) + (long)nritz * j];
    }
    DBuf<int> slots(slots_h.size());
    DBuf<T> hc(hc_h.size()), vd(vec_h.size());
    HIP_TRY(hipMemcpyAsync(slots.p, slots_h.data(), sizeof(int) * slots_h.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(hc.p, hc_h.data(), sizeof(T) * hc_h.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(vd.p, vec_h.data(), sizeof(T) * vec_h.size(), hipMemcpyHostToDevice, st));
    EpiArgs<T> none{};
    hipLaunchKernelGGL((k_u0_first<T>), dim3(g), dim3(RB), 0, st, n, d, b, W.p + (long)n * ind[0], U, vd.p);
    LAUNCH_CHECK();
    for (int i = 1; i <= nritz; i++) {
        const T *vi = W.p + (long)n * ind[i - 1];
        if (M) {
            op_precon<T>(*M, vi, tmp.p, st);
            vi = tmp.p;
        }
        op_apply<T>(A, vi, t.p, w1.p, w2.p, EPI_NONE, none, nullptr, st);
        const int low = std::max(1, i - d);
        const bool last = i == nritz;
        hipLaunchKernelGGL((k_u0_step<T>), dim3(g), dim3(RB), 0, st, n, d, i - low + 1, t.p, W.p,
                           slots.p + (size_t)(i - 1) * (d + 1), hc.p + (size_t)(i - 1) * (d + 2),
                           last ? 0 : ind[i], U, last ? (const T *)nullptr : vd.p + (size_t)i * d);
        LAUNCH_CHECK();
    }
    // U0(:, j) = U0(:, j) / norm(U0(:, j)) (ritz_module.f90:187-192)
    Reducer<T> red(n);
    for (int j = 0; j < d; j++) {
        const double nrm = red.norm(n, U + (long)n * j, st);
        hipLaunchKernelGGL((k_div_real<T>), dim3(g), dim3(RB), 0, st, n, nrm, U + (long)n * j);
        LAUNCH_CHECK();
    }
}

// global_QR_factorization (ritz_module.f90:36-98) in place on Q (n x d,
// device): R(i,j) = inprod(Q_i, Q_j) and Q_j -= R(i,j) Q_i for i < j in
// order, then Q_j /= norm(Q_j).
template <class T>
static void global_qr(long n, int d, T *Q, hipStream_t st) {
    const int g = vec_grid(n);
    Reducer<T> red(n);
    DBuf<T> coef(1);
    for (int j = 0; j < d; j++) {
        T *qj = Q + (long)n * j;
        for (int i = 0; i < j; i++) {
            const T *qi = Q + (long)n * i;
            hipLaunchKernelGGL((k_dot_partials<T>), dim3(red.g), dim3(RB), 0, st, n, qi, qj, red.part.p);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_dot_fin<T>), dim3(1), dim3(FIN_THREADS), 0, st, red.part.p, red.g, coef.p);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_sub_scaled<T>), dim3(g), dim3(RB), 0, st, n, coef.p, qi, qj);
            LAUNCH_CHECK();
        }
        const double r = red.norm(n, qj, st);
        if (!(r > std::numeric_limits<double>::min())) throw std::runtime_error(" V has dependent columns ");
        hipLaunchKernelGGL((k_div_real<T>), dim3(g), dim3(RB), 0, st, n, r, qj);
        LAUNCH_CHECK();
    }
}

template <class T>
int idrs_ritz_solve(const Operator &A, long n, int s, const T *b, T *x, SolveOpts<T> o, RitzRequest<T> &rq,
                    SolveStats &out, hipStream_t st) {
    constexpr bool cplx = Tr<T>::is_complex;
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * n, 1000);   // ritz_module.f90:472-473
    if (maxit < 1) throw std::invalid_argument("IDRS_RITZ needs maximum_iterations >= 1");
    o.has_maxit = true;
    o.maxit = maxit;
    const long ldh = (long)maxit + 1;
    std::vector<T> Hf((size_t)ldh * maxit);
    o.H = Hf.data();
    o.nritz = maxit;
    if (rq.has_M) {
        o.has_m1 = true;
        o.M1 = rq.M;
    }
    if (rq.n_omega < 0 || rq.dimU0 < 0) throw std::invalid_argument("negative n_omega or dimension of U0");
    if (rq.dimU0 > 0 && !rq.U0) throw std::invalid_argument("U0 output missing");
    if (rq.n_omega > 0 && !rq.omega) throw std::invalid_argument("omega output missing");
    if (rq.H && ((long)rq.h_rows * rq.h_cols > ldh * maxit || rq.h_rows > ldh || rq.h_cols > maxit))
        throw std::invalid_argument("Dimensions of Hessenberg matrix too large");   // ritz_module.f90:549
    idrs_solve<T>(A, n, s, b, x, o, out, st);
    const int iter = out.iterations;
    const int nritz = (iter / (s + 1)) * s + iter % (s + 1);                      // ritz_module.f90:486

    std::vector<std::complex<double>> Hs;   // H as std::complex for the complex routines
    if (cplx) {
        Hs.resize(Hf.size());
        for (size_t q = 0; q < Hf.size(); q++) Hs[q] = to_std(Hf[q]);
    }
    // omega (ritz_module.f90:493-521 real, 331-358 complex)
    if (rq.n_omega > 0) {
        if (nritz < rq.n_omega) throw std::invalid_argument("Too many parameters omega requested, reduce omega.");
        if constexpr (cplx) {
            const auto rv = ritz::qr_eig(rq.n_omega, Hs.data(), (int)ldh, nullptr, nullptr);
            const auto om = ritz::ritz_omegas_complex(rv);
            for (int i = 0; i < rq.n_omega; i++) rq.omega[i] = from_std<T>(om[i]);
            rq.converged = rq.n_omega;
        } else {
            int conv = 0;
            const auto rv = ritz::qr_eig(rq.n_omega, (const double *)Hf.data(), (int)ldh, nullptr, &conv);
            const auto om = ritz::ritz_omegas(rv);
            for (int i = 0; i < rq.n_omega; i++) rq.omega[i] = om[i];
            rq.converged = conv;
        }
    }
    // U0 (ritz_module.f90:523-546 real, 360-377 complex)
    if (rq.dimU0 > 0) {
        const int d = rq.dimU0;
        if (nritz < d) throw std::invalid_argument("Recycle space is too large, reduce dimension of U0.");
        std::vector<T> vec((size_t)nritz * d);
        if constexpr (cplx) {
            std::vector<std::complex<double>> V;
            ritz::qr_eig(nritz, Hs.data(), (int)ldh, &V, nullptr);
            for (int j = 0; j < d; j++)
                for (int i = 0; i < nritz; i++) vec[i + (size_t)nritz * j] = from_std<T>(V[i + (size_t)nritz * j]);
        } else {
            const auto sv = ritz::si_power(nritz, (const double *)Hf.data(), (int)ldh, d, 0.0, rq.seed, rq.offset);
            for (size_t q = 0; q < vec.size(); q++) vec[q] = sv[q];
        }
        compute_u0<T>(A, rq.M, n, b, Hf.data(), ldh, nritz, vec.data(), d, rq.U0, st);
        if (rq.M)
            for (int i = 0; i < d; i++) op_precon<T>(*rq.M, rq.U0 + (long)n * i, rq.U0 + (long)n * i, st);
        global_qr<T>(n, d, rq.U0, st);
        HIP_TRY(hipStreamSynchronize(st));
    }
    // H = Hf(1:size(H,1), 1:size(H,2)) (ritz_module.f90:548-551)
    if (rq.H)
        for (int j = 0; j < rq.h_cols; j++)
            for (int i = 0; i < rq.h_rows; i++) rq.H[i + (size_t)rq.h_rows * j] = Hf[i + (size_t)ldh * j];
    return nritz;
}

template int idrs_ritz_solve<double>(const Operator &, long, int, const double *, double *, SolveOpts<double>,
                                     RitzRequest<double> &, SolveStats &, hipStream_t);
template int idrs_ritz_solve<c128>(const Operator &, long, int, const c128 *, c128 *, SolveOpts<c128>,
                                   RitzRequest<c128> &, SolveStats &, hipStream_t);

}  // namespace idrs
