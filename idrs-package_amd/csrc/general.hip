// The generalized operator of matrix_module: mass and damping matrices,
// block systems and deflation (MODULES/PP_MATRIX/matrix_module.f90:169-264
// real_mv/complex_mv, :306-400 rprecon/cprecon, :1284-1623 deflation).
//
// These operators run the reference's statements in order with unscaled
// products of K, M and C (the fast path folds the diagonal scaling into K's
// values, which only covers K - seed I): t = v/P, then the branch's
// products and elementwise combinations, each rounded as the reference
// rounds it.  Deflation's small solves run on the device (one workgroup:
// the V^H w sums, then the reference's Gaussian elimination with partial
// pivoting on V^H AV), so an operator application is a fixed launch
// sequence that HIP graphs can capture.
#include "general.hpp"

#include "lu.hpp"
#include "spmv_epi.hpp"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace idrs {

namespace {

constexpr int GB = 256;   // block size of the elementwise kernels

#define GLOOP(q, n) for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += (long)gridDim.x * blockDim.x)

// w = v/P on the first nn rows (d: P's diagonal, null: copy), a copy beyond
template <class T, class DT>
__global__ __launch_bounds__(GB) void k_gprecon(long nn, long n, const T *__restrict__ v, const DT *__restrict__ d,
                                                T *__restrict__ w) {
    GLOOP(q, n) w[q] = (q < nn && d) ? divv(v[q], d[q]) : v[q];
}

// w = w + a u (sub: w - a u), the product rounded before the sum
template <class T>
__global__ __launch_bounds__(GB) void k_lin(long n, T *__restrict__ w, T a, const T *__restrict__ u, int sub) {
#pragma clang fp contract(off)
    GLOOP(q, n) {
        const T p = a * u[q];
        w[q] = sub ? w[q] - p : w[q] + p;
    }
}

// the block system's shift-and-invert right-hand side (matrix_module.f90:
// 361-366): u = (v1 + u) + s m, m = M v2 (mass) or v2
template <class T>
__global__ __launch_bounds__(GB) void k_si_rhs(long nn, const T *__restrict__ v1, T *__restrict__ u, T s,
                                               const T *__restrict__ m) {
#pragma clang fp contract(off)
    GLOOP(q, nn) {
        const T p = s * m[q];
        u[q] = (v1[q] + u[q]) + p;
    }
}

// w1 = t1 - s t2, w2 = -t2 (t1 = v2)
template <class T>
__global__ __launch_bounds__(GB) void k_si_out(long nn, const T *__restrict__ v2, const T *__restrict__ t2, T s,
                                               T *__restrict__ w) {
#pragma clang fp contract(off)
    GLOOP(q, nn) {
        const T p = s * t2[q];
        const T a = v2[q];
        w[q] = a - p;
        w[nn + q] = -t2[q];
    }
}

// w = (-u) - w
template <class T>
__global__ __launch_bounds__(GB) void k_negsub(long n, const T *__restrict__ u, T *__restrict__ w) {
#pragma clang fp contract(off)
    GLOOP(q, n) w[q] = (-u[q]) - w[q];
}

template <class T>
__global__ __launch_bounds__(GB) void k_gcopy(long n, const T *__restrict__ src, T *__restrict__ dst) {
    GLOOP(q, n) dst[q] = src[q];
}

// w = a v
template <class T>
__global__ __launch_bounds__(GB) void k_gscale(long n, T a, const T *__restrict__ v, T *__restrict__ w) {
    GLOOP(q, n) w[q] = a * v[q];
}

// w = t + a v (polpre's step, the product rounded before the sum)
template <class T>
__global__ __launch_bounds__(GB) void k_gaxpy(long n, const T *__restrict__ t, T a, const T *__restrict__ v,
                                              T *__restrict__ w) {
#pragma clang fp contract(off)
    GLOOP(q, n) {
        const T p = a * v[q];
        w[q] = t[q] + p;
    }
}

// w = x1 + x2
template <class T>
__global__ __launch_bounds__(GB) void k_gadd(long n, const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ w) {
    GLOOP(q, n) w[q] = a[q] + b[q];
}

// block partials of s(j) = sum_i conj(V(i,j)) w(i), j < nd
template <class T>
__global__ __launch_bounds__(GB) void k_vdots(long n, int nd, const T *__restrict__ V, const T *__restrict__ w,
                                              T *partials) {
    __shared__ T lds[MAX_DEFL * (GB / WAVE)];
    T e[MAX_DEFL];
#pragma unroll
    for (int j = 0; j < MAX_DEFL; j++) e[j] = T(0);
    GLOOP(q, n) {
        const T wq = w[q];
#pragma unroll
        for (int j = 0; j < MAX_DEFL; j++)
            if (j < nd) e[j] = e[j] + Tr<T>::conj(V[q + (long)n * j]) * wq;
    }
    block_sum<T, MAX_DEFL>(e, lds);
    if (threadIdx.x == 0)
        for (int j = 0; j < nd; j++) partials[(long)j * gridDim.x + blockIdx.x] = e[j];
}

// SOLVE (dense_la_module.f90:605-681): Gaussian elimination with partial
// pivoting on the augmented matrix, one thread, in the reference's order
template <class T>
__device__ inline void dev_solve(int n, T *M /* n x (n+1), column-major */, T *x) {
#pragma clang fp contract(off)
    auto at = [&](int i, int c) -> T & { return M[i + n * c]; };
    for (int j = 0; j < n - 1; j++) {
        int k = j;
        double best = Tr<T>::absv(at(j, j));
        for (int i = j + 1; i < n; i++)
            if (Tr<T>::absv(at(i, j)) > best) {
                best = Tr<T>::absv(at(i, j));
                k = i;
            }
        if (k != j)
            for (int c = 0; c <= n; c++) {
                const T tmp = at(k, c);
                at(k, c) = at(j, c);
                at(j, c) = tmp;
            }
        for (int i = j + 1; i < n; i++) {
            const T p = at(i, j) / at(j, j);
            for (int c = j + 1; c <= n; c++) at(i, c) = at(i, c) - at(j, c) * p;
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        T p = at(i, n);
        for (int j = i + 1; j < n; j++) p = p - at(i, j) * at(j, n);
        at(i, n) = p / at(i, i);
    }
    for (int i = 0; i < n; i++) x[i] = at(i, n);
}

// t = the sums s (solve: t = solve(E - shift I, s), E = V^H AV)
template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_defl_fin(const T *partials, int nparts, int nd, const T *E, T shift,
                                                          int solve, T *t) {
#pragma clang fp contract(off)
    __shared__ T red[MAX_DEFL];
    reduce_partials(partials, nparts, nd, red);
    if (threadIdx.x != 0) return;
    if (!solve) {
        for (int j = 0; j < nd; j++) t[j] = red[j];
        return;
    }
    T M[MAX_DEFL * (MAX_DEFL + 1)];
    for (int c = 0; c < nd; c++)
        for (int i = 0; i < nd; i++) M[i + nd * c] = E[i + nd * c];
    for (int i = 0; i < nd; i++) M[i + nd * i] = M[i + nd * i] - shift;
    for (int i = 0; i < nd; i++) M[i + nd * nd] = red[i];
    dev_solve(nd, M, t);
}

// w = w - matmul(V, t) (sign -1) or w = matmul(V, t) (sign 0); each row's
// sum over j in order from zero
template <class T>
__global__ __launch_bounds__(GB) void k_defl_sub(long n, int nd, const T *__restrict__ V, const T *__restrict__ t,
                                                 T *__restrict__ w, int sub) {
#pragma clang fp contract(off)
    __shared__ T st[MAX_DEFL];
    if (threadIdx.x < nd) st[threadIdx.x] = t[threadIdx.x];
    __syncthreads();
    GLOOP(q, n) {
        T acc = T(0);
        for (int j = 0; j < nd; j++) acc = acc + V[q + (long)n * j] * st[j];
        w[q] = sub ? w[q] - acc : acc;
    }
}

// the SpMV epilogues' block partials of an operator application that ended
// outside an SpMV kernel (E.nparts workgroups)
template <class T, int EPI, int NV>
__global__ __launch_bounds__(SPMV_THREADS) void k_epi(long n, const T *__restrict__ y, EpiArgs<T> E, const int *info) {
    if (info && *info >= 0) return;
    constexpr int PC = EPI == EPI_PDOT ? NV : 1;
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    GLOOP(q, n) {
        T pe[PC];
        epi_load<T, EPI, PC>(pe, E, q, 0, 1);
        epi_accum<T, EPI, NV, PC>(e, pe, E, y[q], 0, 1);
    }
    epi_store<T, EPI, NV>(e, E, s_red);
}

template <class T>
T seedT(const Operator &op) {
    if constexpr (Tr<T>::is_complex) return T(op.seed.real(), op.seed.imag());
    else return op.seed.real();
}

template <class T>
T fromc(std::complex<double> v) {
    if constexpr (Tr<T>::is_complex) return T(v.real(), v.imag());
    else return v.real();
}

// workspace of the general path: eight vectors of rows() elements of T
// (6 and 7 private to the preconditioner), then the deflation partials and
// coefficients
struct GWork {
    char *base;
    long rows;
    template <class T> T *vec(int i) const { return (T *)base + (long)i * rows; }
    template <class T> T *part() const { return (T *)base + 8L * rows; }
    template <class T> T *coef() const { return part<T>() + (long)MAX_DEFL * GEN_PARTS; }
};

template <class T>
GWork work(const Operator &op) {
    if (!op.gws) throw std::invalid_argument("operator workspace missing: prepare the operator first");
    return GWork{(char *)op.gws, op.rows()};
}

int ggrid(long n) { return std::max(1, std::min(vec_grid(n), GEN_PARTS)); }

template <class T>
void lin(long n, T *w, T a, const T *u, int sub, hipStream_t st) {
    hipLaunchKernelGGL((k_lin<T>), dim3(ggrid(n)), dim3(GB), 0, st, n, w, a, u, sub);
    LAUNCH_CHECK();
}

template <class T>
void gcopy(long n, const T *src, T *dst, hipStream_t st) {
    hipLaunchKernelGGL((k_gcopy<T>), dim3(ggrid(n)), dim3(GB), 0, st, n, src, dst);
    LAUNCH_CHECK();
}

// t = coefficients of the deflation step for w: V^H w, solved with
// E - shift I when solve is set
template <class T>
void defl_coef(const Operator &op, const T *w, T shift, int solve, T *t, hipStream_t st) {
    const GWork g = work<T>(op);
    const long n = op.rows();
    const int gp = ggrid(n);
    hipLaunchKernelGGL((k_vdots<T>), dim3(gp), dim3(GB), 0, st, n, op.ndefl, (const T *)op.dV, w, g.part<T>());
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_defl_fin<T>), dim3(1), dim3(FIN_THREADS), 0, st, g.part<T>(), gp, op.ndefl,
                       (const T *)op.dVtAV, shift, solve, t);
    LAUNCH_CHECK();
}

template <class T>
void defl_sub(const Operator &op, const T *V, const T *t, T *w, int sub, hipStream_t st) {
    const long n = op.rows();
    hipLaunchKernelGGL((k_defl_sub<T>), dim3(ggrid(n)), dim3(GB), 0, st, n, op.ndefl, V, t, w, sub);
    LAUNCH_CHECK();
}

}  // namespace

// v/A%P on nn rows (ruser_precon / cuser_precon, user_module.f90:243-283):
// the LU solve, the diagonal, or a copy
template <class T>
static void user_precon(const Operator &op, long nn, const T *v, T *w, hipStream_t st) {
    if (op.lu) {
        lu_solve<T>(*op.lu, v, w, st);
        return;
    }
    const int g = ggrid(nn);
    const bool d = op.which_precon == 1 && op.pdiag;
    if (d && op.pdiag_kind == ValKind::Complex) {
        if constexpr (Tr<T>::is_complex)
            hipLaunchKernelGGL((k_gprecon<T, c128>), dim3(g), dim3(GB), 0, st, nn, nn, v, (const c128 *)op.pdiag, w);
        else
            throw std::invalid_argument("complex preconditioner for a real solve");
    } else {
        hipLaunchKernelGGL((k_gprecon<T, double>), dim3(g), dim3(GB), 0, st, nn, nn, v,
                           d ? (const double *)op.pdiag : (const double *)nullptr, w);
    }
    LAUNCH_CHECK();
}

// v/A (rprecon/cprecon, matrix_module.f90:306-400): v/A%P; a block system
// applies it to the first half and copies the second, or with
// shift-and-invert inverts [-C-sM -K; I -sI] through P = K + sC + s^2 M:
// t2 = (v1 + C v2 + s (M v2 | v2))/P, w = (v2 - s t2, -t2)
template <class T>
void gen_precon(const Operator &op, const T *v, T *w, hipStream_t st) {
    const long nn = op.K->nrows, n = op.rows();
    if (op.block && op.shift_invert) {
        if (!op.Cm) throw std::invalid_argument("a block system needs the damping matrix C");
        const GWork g = work<T>(op);
        T *u = g.vec<T>(6), *m = g.vec<T>(7);
        const T s = fromc<T>(op.si_shift);
        user_spmv<T>(*op.Cm, v + nn, u, st);
        if (op.Mm) user_spmv<T>(*op.Mm, v + nn, m, st);
        hipLaunchKernelGGL((k_si_rhs<T>), dim3(ggrid(nn)), dim3(GB), 0, st, nn, v, u, s, op.Mm ? (const T *)m : v + nn);
        LAUNCH_CHECK();
        user_precon<T>(op, nn, u, m, st);
        hipLaunchKernelGGL((k_si_out<T>), dim3(ggrid(nn)), dim3(GB), 0, st, nn, v + nn, (const T *)m, s, w);
        LAUNCH_CHECK();
        return;
    }
    user_precon<T>(op, nn, v, w, st);
    if (n > nn) gcopy<T>(n - nn, v + nn, w + nn, st);
}

// real_mv/complex_mv (matrix_module.f90:169-264); t, u: workspace vectors
template <class T>
void gen_mv(const Operator &op, const T *v, T *w, T *t, T *u, hipStream_t st) {
    const long nn = op.K->nrows, n = op.rows();
    const T seed = seedT<T>(op);
    const bool has_seed = op.seed != std::complex<double>(0.0, 0.0);
    gen_precon<T>(op, v, t, st);
    if (op.block) {                                              // :189-201
        if (!op.Cm) throw std::invalid_argument("a block system needs the damping matrix C");
        user_spmv<T>(*op.Cm, t, u, st);
        user_spmv<T>(*op.K, t + nn, w, st);
        hipLaunchKernelGGL((k_negsub<T>), dim3(ggrid(nn)), dim3(GB), 0, st, nn, u, w);
        LAUNCH_CHECK();
        gcopy<T>(nn, t, w + nn, st);
        if (has_seed) {
            if (op.Mm) {
                user_spmv<T>(*op.Mm, t, u, st);
                lin<T>(nn, w, seed, u, 1, st);
                lin<T>(nn, w + nn, seed, t + nn, 1, st);
            } else {
                lin<T>(n, w, seed, t, 1, st);
            }
        }
        return;
    }
    user_spmv<T>(*op.K, t, w, st);                              // :203-214
    if (!has_seed) return;
    const T s2 = seed * seed;
    if (op.Cm && op.Mm) {
        user_spmv<T>(*op.Cm, t, u, st);
        lin<T>(n, w, seed, u, 0, st);
        user_spmv<T>(*op.Mm, t, u, st);
        lin<T>(n, w, s2, u, 0, st);
    } else if (op.Cm) {
        user_spmv<T>(*op.Cm, t, u, st);
        lin<T>(n, w, seed, u, 0, st);
        lin<T>(n, w, s2, t, 0, st);
    } else if (op.Mm) {
        user_spmv<T>(*op.Mm, t, u, st);
        lin<T>(n, w, seed, u, 1, st);
    } else {
        lin<T>(n, w, seed, t, 1, st);
    }
}

// Pmul (matrix_module.f90:1369-1431) in place
template <class T>
void gen_pmul(const Operator &op, T *w, hipStream_t st) {
    if (op.ndefl <= 0) return;
    const GWork g = work<T>(op);
    T *t = g.coef<T>();
    if (op.multishift) {
        defl_coef<T>(op, w, T(0), 0, t, st);
        defl_sub<T>(op, (const T *)op.dV, t, w, 1, st);
    } else {
        defl_coef<T>(op, w, T(0), 1, t, st);
        // the real version subtracts matmul(A%rV, t) (:1415), the complex
        // one matmul(A%cAV, t) (:1384)
        defl_sub<T>(op, Tr<T>::is_complex ? (const T *)op.dAV : (const T *)op.dV, t, w, 1, st);
    }
}

// polpre (matrix_module.f90:266-304) with gen_mv, into w
template <class T>
static void gen_polpre(const Operator &op, const std::vector<std::complex<double>> &alpha, int d, const T *v, T *w,
                       hipStream_t st) {
    const GWork g = work<T>(op);
    const long n = op.rows();
    const int gr = ggrid(n);
    if (d == 0) {
        gcopy<T>(n, v, w, st);
        return;
    }
    T *tmp = g.vec<T>(1), *t = g.vec<T>(2), *u = g.vec<T>(3);
    hipLaunchKernelGGL((k_gscale<T>), dim3(gr), dim3(GB), 0, st, n, fromc<T>(alpha[d]), v, w);
    LAUNCH_CHECK();
    for (int i = d - 1; i >= 0; i--) {
        gen_mv<T>(op, w, tmp, t, u, st);
        hipLaunchKernelGGL((k_gaxpy<T>), dim3(gr), dim3(GB), 0, st, n, tmp, fromc<T>(alpha[i]), v, w);
        LAUNCH_CHECK();
    }
}

// rmatvec/cmatvec (matrix_module.f90:135-167): w = Pmul(real_mv(polpre(v))),
// then the SpMV epilogue's partials of w
template <class T>
void gen_apply(const Operator &op, const T *v, T *w, int epi, const EpiArgs<T> &E, const int *info, hipStream_t st) {
    const GWork g = work<T>(op);
    const long n = op.rows();
    T *p = g.vec<T>(0), *t = g.vec<T>(2), *u = g.vec<T>(3);
    gen_polpre<T>(op, op.alpha, op.degree, v, p, st);
    gen_mv<T>(op, p, w, t, u, st);
    gen_pmul<T>(op, w, st);
    if (epi == EPI_NONE) return;
    // as many block partials as the SpMV over K would have written: the
    // solver's reductions read spmv_grid(K) of them
    EpiArgs<T> Eg = E;
    Eg.nparts = spmv_grid(*op.K);
    const dim3 gd(Eg.nparts), bd(SPMV_THREADS);
    if (epi == EPI_OMEGA) hipLaunchKernelGGL((k_epi<T, EPI_OMEGA, 3>), gd, bd, 0, st, n, w, Eg, info);
    else if (epi == EPI_NORM) hipLaunchKernelGGL((k_epi<T, EPI_NORM, 1>), gd, bd, 0, st, n, w, Eg, info);
    else {
        switch (Eg.prow) {
#define EPI_CASE(NV) \
    case NV: hipLaunchKernelGGL((k_epi<T, EPI_PDOT, NV>), gd, bd, 0, st, n, w, Eg, info); break;
            EPI_CASE(1) EPI_CASE(2) EPI_CASE(3) EPI_CASE(4) EPI_CASE(5) EPI_CASE(6) EPI_CASE(7) EPI_CASE(8)
            EPI_CASE(9) EPI_CASE(10) EPI_CASE(11) EPI_CASE(12) EPI_CASE(13) EPI_CASE(14) EPI_CASE(15) EPI_CASE(16)
#undef EPI_CASE
            default: throw std::invalid_argument("P^T y epilogue over more than 16 rows");
        }
    }
    LAUNCH_CHECK();
}

// make_deflation (matrix_module.f90:1284-1352): V (device, rows x nd) kept,
// AV = real_mv(A, V), V^H AV
template <class T>
void gen_make_deflation(Operator &op, const T *V, int nd, hipStream_t st) {
    if (nd < 1 || nd > MAX_DEFL) throw std::invalid_argument("deflation takes 1..16 vectors");
    const long n = op.rows();
    (void)hipFree(op.dV);
    (void)hipFree(op.dAV);
    (void)hipFree(op.dVtAV);
    op.dV = op.dAV = op.dVtAV = nullptr;
    op.ndefl = 0;
    HIP_TRY(hipMalloc(&op.dV, sizeof(T) * n * nd));
    HIP_TRY(hipMalloc(&op.dAV, sizeof(T) * n * nd));
    HIP_TRY(hipMalloc(&op.dVtAV, sizeof(T) * nd * nd));
    op.defl_complex = Tr<T>::is_complex;
    HIP_TRY(hipMemcpyAsync(op.dV, V, sizeof(T) * n * nd, hipMemcpyDeviceToDevice, st));
    op.ndefl = nd;
    gen_prepare(op);
    const GWork g = work<T>(op);
    T *t = g.vec<T>(2), *u = g.vec<T>(3);
    // AV = real_mv(A, V) (no polynomial, no Pmul)
    for (int j = 0; j < nd; j++) gen_mv<T>(op, (const T *)op.dV + (long)n * j, (T *)op.dAV + (long)n * j, t, u, st);
    // VtAV(i, j) = dot_product(V(:,i), AV(:,j)): the sums of column j of AV
    // against every V column, by the same kernel as Pmul's
    T *E = (T *)op.dVtAV;
    for (int j = 0; j < nd; j++) defl_coef<T>(op, (const T *)op.dAV + (long)n * j, T(0), 0, E + (long)nd * j, st);
    HIP_TRY(hipStreamSynchronize(st));
}

// add_deflation (matrix_module.f90:1433-1506) / rms_add_deflation (one
// shift, sigma_s = sigma - seed; :1508-1623): x = V Hs^-1 V^H b +
// (y - V Hs^-1 V^H (real_mv(A, y) - sigma_s y)), Hs = V^H AV - sigma_s I
template <class T>
static void add_defl(const Operator &op, const T *b, const T *y, T *x, bool shifted, T ss, hipStream_t st) {
    const long n = op.rows();
    if (op.ndefl <= 0) {
        gcopy<T>(n, y, x, st);
        return;
    }
    const GWork g = work<T>(op);
    T *x1 = g.vec<T>(4), *Ay = g.vec<T>(5), *t = g.vec<T>(2), *u = g.vec<T>(3), *c = g.coef<T>();
    const T *V = (const T *)op.dV;
    defl_coef<T>(op, b, shifted ? ss : T(0), 1, c, st);
    defl_sub<T>(op, V, c, x1, 0, st);                      // x1 = matmul(V, t)
    gen_mv<T>(op, y, Ay, t, u, st);
    if (shifted) lin<T>(n, Ay, ss, y, 1, st);
    defl_coef<T>(op, Ay, shifted ? ss : T(0), 1, c, st);
    gcopy<T>(n, y, x, st);
    defl_sub<T>(op, V, c, x, 1, st);                       // x2 = y - matmul(V, t)
    hipLaunchKernelGGL((k_gadd<T>), dim3(ggrid(n)), dim3(GB), 0, st, n, x1, x, x);
    LAUNCH_CHECK();
}

template <class T>
void gen_add_deflation(const Operator &op, const T *b, const T *y, T *x, hipStream_t st) {
    add_defl<T>(op, b, y, x, false, T(0), st);
}

// sigma_s of a shift: (sigma - seed)/(sigma - shift) with shift-and-invert,
// else sigma - seed (matrix_module.f90:443-456, 1534-1538)
template <class T>
static T sigma_s(const Operator &op, std::complex<double> sigma) {
    const T num = fromc<T>(sigma) - seedT<T>(op);
    if (!op.shift_invert) return num;
    const T den = fromc<T>(sigma) - fromc<T>(op.si_shift);
    return num / den;
}

template <class T>
void gen_ms_add_deflation(const Operator &op, int nsig, const std::complex<double> *sigma, const T *b, const T *Y,
                          T *X, hipStream_t st) {
    const long n = op.rows();
    for (int k = 0; k < nsig; k++) {
        const T ss = sigma_s<T>(op, sigma[k]);
        add_defl<T>(op, b, Y + (long)n * k, X + (long)n * k, true, ss, st);
    }
}

// scaleback (matrix_module.f90:808-849): x = polpre(y); x = add_deflation(b, x); x = x/A
template <class T>
void gen_scaleback(const Operator &op, const T *b, const T *y, T *x, hipStream_t st) {
    const GWork g = work<T>(op);
    T *p = g.vec<T>(0), *q = g.vec<T>(1);
    gen_polpre<T>(op, op.alpha, op.degree, y, p, st);
    // add_defl uses vectors 2..5
    add_defl<T>(op, b, p, q, false, T(0), st);
    gen_precon<T>(op, q, x, st);
}

// ms_rscaleback/ms_cscaleback (matrix_module.f90:724-806) for one shift's
// column: x = p_sigma(A) y with the shift's coefficients, the deflation
// step for that shift, the shift-and-invert factor
// (seed - shift)/(sigma - shift), x/A
template <class T>
void gen_ms_scaleback_col(const Operator &op, const std::vector<std::complex<double>> &alpha, std::complex<double> sigma,
                          const T *b, const T *y, T *x, hipStream_t st) {
    const GWork g = work<T>(op);
    T *p = g.vec<T>(0), *q = g.vec<T>(1);
    gen_polpre<T>(op, alpha, (int)alpha.size() - 1, y, p, st);
    if (op.ndefl > 0) {
        add_defl<T>(op, b, p, q, true, sigma_s<T>(op, sigma), st);
    } else {
        gcopy<T>(op.rows(), p, q, st);
    }
    if (op.shift_invert) {
        const T sh = fromc<T>(op.si_shift);
        const T f = (seedT<T>(op) - sh) / (fromc<T>(sigma) - sh);
        hipLaunchKernelGGL((k_gscale<T>), dim3(ggrid(op.rows())), dim3(GB), 0, st, op.rows(), f, (const T *)q, p);
        LAUNCH_CHECK();
        gen_precon<T>(op, p, x, st);
        return;
    }
    gen_precon<T>(op, q, x, st);
}

void gen_prepare(const Operator &op) {
    if (!op.general() || (op.gws && op.gws_rows == op.rows())) return;
    (void)hipFree(op.gws);
    op.gws = nullptr;
    const long n = op.rows();
    HIP_TRY(hipMalloc(&op.gws, sizeof(c128) * (8 * n + (long)MAX_DEFL * GEN_PARTS + MAX_DEFL)));
    op.gws_rows = n;
}

#define GINST(T)                                                                                                 \
    template void gen_precon<T>(const Operator &, const T *, T *, hipStream_t);                                   \
    template void gen_mv<T>(const Operator &, const T *, T *, T *, T *, hipStream_t);                             \
    template void gen_pmul<T>(const Operator &, T *, hipStream_t);                                                \
    template void gen_apply<T>(const Operator &, const T *, T *, int, const EpiArgs<T> &, const int *, hipStream_t); \
    template void gen_make_deflation<T>(Operator &, const T *, int, hipStream_t);                                 \
    template void gen_add_deflation<T>(const Operator &, const T *, const T *, T *, hipStream_t);                 \
    template void gen_ms_add_deflation<T>(const Operator &, int, const std::complex<double> *, const T *,         \
                                          const T *, T *, hipStream_t);                                           \
    template void gen_scaleback<T>(const Operator &, const T *, const T *, T *, hipStream_t);                     \
    template void gen_ms_scaleback_col<T>(const Operator &, const std::vector<std::complex<double>> &,            \
                                          std::complex<double>, const T *, const T *, T *, hipStream_t);
GINST(double)
GINST(c128)
#undef GINST

}  // namespace idrs
