// Shift-and-invert LU preconditioners on the device (the reference's
// user_module.f90:1383-1459 assembly, :1604-1701 / :2068-2179 factorization
// and :1703-1831 / :2181-2336 solves).
//
// Layout: row-major, so a row of the band (2 hb + 1 entries) or of the dense
// matrix is contiguous for the row-oriented solves.  Factorization keeps the
// reference's elimination order and roundings (l = a_ik / pivot rounded, then
// a_ic - l u_kc with the product rounded before the difference), so the
// factors equal the oracle's; the solves form each row's sum as a block
// reduction, the only departure from the reference's column-by-column order.
//
// Band factorization: one workgroup walks the pivots; at pivot k its threads
// update the hb x hb trailing entries (rows k+1..k+hb, columns k+1..k+hb),
// then divide the hb multipliers.  Dense factorization: two launches per
// pivot over the trailing matrix.  Solves: one workgroup, one row per step
// (forward with the unit lower factor, then back substitution).  These are
// sequential-pivot algorithms; they run where the reference runs them, at
// set-up and once per preconditioner application.
#include "lu.hpp"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace idrs {

LuFactor::~LuFactor() { (void)hipFree(lu); }

namespace {

constexpr int LB = 256;          // elementwise kernels
constexpr int FACT_THREADS = 1024;

#define LLOOP(q, n) for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += (long)gridDim.x * blockDim.x)

__device__ __host__ inline long lu_ix(int dense, long ld, int hb, long i, long c) {
    return dense ? i * ld + c : i * ld + (c - i + hb);
}

template <class FT>
__device__ inline FT val_of(int kind, const void *val, long p) {
    if constexpr (Tr<FT>::is_complex) {
        if (kind == (int)ValKind::Complex) return ((const c128 *)val)[p];
        return FT(kind == (int)ValKind::Real ? ((const double *)val)[p] : 1.0, 0.0);
    } else {
        return kind == (int)ValKind::Real ? ((const double *)val)[p] : 1.0;
    }
}

// real_dense / real_band of one matrix (entries of a row summed in storage
// order: CRS order, COO file order after the stable row sort)
template <class FT>
__global__ __launch_bounds__(LB) void k_scatter_crs(long n, int dense, long ld, int hb, const int *__restrict__ rp,
                                                    const int *__restrict__ col, int kind, const void *val, FT *out) {
    LLOOP(i, n) {
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            const long q = lu_ix(dense, ld, hb, i, col[p]);
            out[q] = out[q] + val_of<FT>(kind, val, p);
        }
    }
}

template <class FT>
__global__ __launch_bounds__(LB) void k_scatter_dia(long n, long ncols, long ld, int hb, int ndiag,
                                                    const int *__restrict__ off, int kind, const void *val, FT *out) {
    LLOOP(i, n) {
        for (int d = 0; d < ndiag; d++) {
            const long c = i + off[d];
            if (c < 0 || c >= ncols) continue;
            const long q = lu_ix(0, ld, hb, i, c);
            out[q] = out[q] + val_of<FT>(kind, val, (long)d * n + i);
        }
    }
}

template <class FT>
__global__ __launch_bounds__(LB) void k_zero(long n, FT *a) {
    LLOOP(q, n) a[q] = FT(0);
}

// the shifted combination (user_module.f90:1392-1419, :1427-1449):
// mode 0: K + s C + s^2 M, 1: K + s C, 2: K - s M
template <class FT>
__global__ __launch_bounds__(LB) void k_combine(long cnt, int mode, FT s, FT s2, FT *__restrict__ a,
                                                const FT *__restrict__ c, const FT *__restrict__ m) {
#pragma clang fp contract(off)
    LLOOP(q, cnt) {
        if (mode == 0) {
            const FT t1 = s * c[q], t2 = s2 * m[q];
            a[q] = (a[q] + t1) + t2;
        } else if (mode == 1) {
            const FT t1 = s * c[q];
            a[q] = a[q] + t1;
        } else {
            const FT t2 = s * m[q];
            a[q] = a[q] - t2;
        }
    }
}

// a(i,i) = a(i,i) + d
template <class FT>
__global__ __launch_bounds__(LB) void k_add_diag(long n, int dense, long ld, int hb, FT d, FT *a) {
    LLOOP(i, n) {
        const long q = lu_ix(dense, ld, hb, i, i);
        a[q] = a[q] + d;
    }
}

// band_LU_decomp: one workgroup; info = k + 1 at a zero pivot k
template <class FT>
__global__ __launch_bounds__(FACT_THREADS) void k_band_lu(long n, int hb, long ld, FT *B, int *info) {
#pragma clang fp contract(off)
    for (long k = 0; k < n - 1; k++) {
        const FT piv = B[k * ld + hb];
        if (piv == FT(0)) {
            if (threadIdx.x == 0) *info = (int)(k + 1);
            return;
        }
        const long rows = std::min(n - 1, k + hb) - k;
        const long pairs = rows * hb;
        // columns k+1..k+hb of rows k+1..k+rows (padding past n included, as
        // the reference updates the whole band row); multipliers unchanged yet
        for (long p = threadIdx.x; p < pairs; p += blockDim.x) {
            const long i = k + 1 + p / hb;
            const int j = 1 + (int)(p % hb);
            const FT l = B[i * ld + (k - i + hb)] / piv;
            const FT prod = l * B[k * ld + hb + j];
            FT &e = B[i * ld + (k + j - i + hb)];
            e = e - prod;
        }
        __syncthreads();
        for (long r = threadIdx.x; r < rows; r += blockDim.x) {
            const long i = k + 1 + r;
            B[i * ld + (k - i + hb)] = B[i * ld + (k - i + hb)] / piv;
        }
        __syncthreads();
    }
}

// RLU_decomp at pivot k, trailing update: rows i > k with a(i,k) /= 0,
// columns c > k (the multiplier rounded as the reference rounds it)
template <class FT>
__global__ __launch_bounds__(LB) void k_dense_update(long n, long k, FT *A, int *info) {
#pragma clang fp contract(off)
    if (*info) return;
    const FT piv = A[k * n + k];
    if (piv == FT(0)) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *info = (int)(k + 1);
        return;
    }
    const long w = n - k - 1;
    const long c = k + 1 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (long i = k + 1 + blockIdx.y; i < n; i += gridDim.y) {
        const FT a = A[i * n + k];
        if (a == FT(0)) continue;
        const FT l = a / piv;
        const FT prod = l * A[k * n + c];
        A[i * n + c] = A[i * n + c] - prod;
    }
    (void)w;
}

// ... then the multipliers of column k
template <class FT>
__global__ __launch_bounds__(LB) void k_dense_lcol(long n, long k, FT *A, const int *info) {
    if (*info) return;
    const FT piv = A[k * n + k];
    for (long i = k + 1 + (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const FT a = A[i * n + k];
        if (a != FT(0)) A[i * n + k] = a / piv;
    }
}

template <class T, class FT>
__device__ inline T mulv(FT a, T b) {
    if constexpr (Tr<T>::is_complex && !Tr<FT>::is_complex) return a * b;
    else return a * b;
}

// x = (LU)^-1 b, one workgroup of NT threads, row by row: forward
// x(i) = b(i) - sum_c L(i,c) x(c), then x(i) = (x(i) - sum_c U(i,c) x(c)) / U(i,i).
// x is written by thread 0 and read by the workgroup after the barrier.
template <class T, class FT, int NT>
__global__ __launch_bounds__(NT) void k_lu_solve(long n, int dense, int hb, long ld, const FT *__restrict__ F,
                                                 const T *__restrict__ b, T *x) {
#pragma clang fp contract(off)
    __shared__ T red[NT / WAVE];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    auto reduce = [&](T acc) -> T {
        acc = wave_sum(acc);
        if constexpr (NT == WAVE) {
            return acc;
        } else {
            if (lane == 0) red[wid] = acc;
            __syncthreads();
            T s = red[0];
            for (int w = 1; w < NT / WAVE; w++) s = s + red[w];
            return s;
        }
    };
    for (long i = 0; i < n; i++) {
        const long lo = dense ? 0 : std::max(0L, i - hb);
        T acc = T(0);
        for (long c = lo + threadIdx.x; c < i; c += NT) acc = acc + mulv<T, FT>(F[lu_ix(dense, ld, hb, i, c)], x[c]);
        const T s = reduce(acc);
        if (threadIdx.x == 0) x[i] = b[i] - s;
        __syncthreads();
    }
    for (long i = n - 1; i >= 0; i--) {
        const long hi = dense ? n - 1 : std::min(n - 1, i + hb);
        T acc = T(0);
        for (long c = i + 1 + threadIdx.x; c <= hi; c += NT) acc = acc + mulv<T, FT>(F[lu_ix(dense, ld, hb, i, c)], x[c]);
        const T s = reduce(acc);
        if (threadIdx.x == 0) {
            const T r = x[i] - s;
            x[i] = r / F[lu_ix(dense, ld, hb, i, i)];
        }
        __syncthreads();
    }
}

// Wavefront helpers on doubles: lane 0's value in every lane (read first
// lane: the whole wavefront is active), and a shift of the whole wavefront
// by one lane towards lane 0 (DPP wave_shl:1, lane 63 receives 0).
__device__ inline double lane0(double v) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ inline c128 lane0(c128 v) { return {lane0(v.re), lane0(v.im)}; }
__device__ inline double shl1(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ inline c128 shl1(c128 v) { return {shl1(v.re), shl1(v.im)}; }

// Band solve for a half bandwidth below 64 in the reference's column order
// (band_LU_solve, user_module.f90:2181-2258), one wavefront: lane j holds
// the running value of row i+j (forward) or i-j (back).  A step finishes
// lane 0's row, gives it to every lane (read first lane), subtracts its
// multiple of the pivot column in lanes 1..hb, and shifts the window by
// one lane (DPP), a new row entering at lane hb.  Each row receives its
// subtractions in the reference's order, each product rounded before the
// difference, so the result equals the column sweep's bit for bit.  The
// columns and entering values are loaded PF steps ahead.
template <class T, class FT>
__global__ __launch_bounds__(WAVE) void k_band_solve_wave(long n, int hb, long ld, const FT *__restrict__ F,
                                                          const T *__restrict__ b, T *x) {
#pragma clang fp contract(off)
    constexpr int PF = 16;
    const int j = threadIdx.x;
    const bool act = j >= 1 && j <= hb;
    // forward: column i of L is L(i+j, i) at (i+j)*ld + hb - j; b(i+1+hb) enters after step i
    auto fcol = [&](long i) -> FT { return (act && i + j < n) ? F[(i + j) * ld + hb - j] : FT(0); };
    auto fin = [&](long i) -> T { return (j == hb && i + 1 + hb < n) ? b[i + 1 + hb] : T(0); };
    FT cq[PF];
    T eq[PF];
#pragma unroll
    for (int t = 0; t < PF; t++) {
        cq[t] = fcol(t);
        eq[t] = fin(t);
    }
    T p = (j <= hb && j < n) ? b[j] : T(0);
    for (long i0 = 0; i0 < n; i0 += PF) {
#pragma unroll
        for (int t = 0; t < PF; t++) {
            const long i = i0 + t;
            if (i < n) {
                const T xi = lane0(p);
                if (j == 0) x[i] = xi;
                const T prod = xi * cq[t];
                p = p - prod;
                p = shl1(p);
                if (j == hb) p = eq[t];
                cq[t] = fcol(i + PF);
                eq[t] = fin(i + PF);
            }
        }
    }
    __syncthreads();
    // back: column k of U is U(k-j, k) at (k-j)*ld + hb + j, U(k,k) at k*ld + hb;
    // y(k-1-hb) enters after step k
    auto bcol = [&](long k) -> FT { return (act && k >= 0 && k - j >= 0) ? F[(k - j) * ld + hb + j] : FT(0); };
    auto bdia = [&](long k) -> FT { return k >= 0 ? F[k * ld + hb] : FT(1); };
    auto bin = [&](long k) -> T { return (j == hb && k - 1 - hb >= 0) ? x[k - 1 - hb] : T(0); };
    FT dq[PF];
#pragma unroll
    for (int t = 0; t < PF; t++) {
        cq[t] = bcol(n - 1 - t);
        dq[t] = bdia(n - 1 - t);
        eq[t] = bin(n - 1 - t);
    }
    p = (j <= hb && n - 1 - j >= 0) ? x[n - 1 - j] : T(0);
    for (long k0 = n - 1; k0 >= 0; k0 -= PF) {
#pragma unroll
        for (int t = 0; t < PF; t++) {
            const long k = k0 - t;
            if (k >= 0) {
                const T pk = lane0(p);
                const T xk = pk / dq[t];
                if (j == 0) x[k] = xk;
                const T prod = xk * cq[t];
                p = p - prod;
                p = shl1(p);
                if (j == hb) p = eq[t];
                cq[t] = bcol(k - PF);
                dq[t] = bdia(k - PF);
                eq[t] = bin(k - PF);
            }
        }
    }
}

int egrid(long n) { return (int)std::max(1L, std::min((n + LB - 1) / LB, 4096L)); }

// get_band (user_module.f90:1833-1872) of a device matrix
int band_of(const DevMatrix &A, hipStream_t st) {
    if (A.fmt == DevFmt::DIA) {
        int hb = 0;
        for (int o : A.h_offsets) hb = std::max(hb, std::abs(o));
        return hb;
    }
    std::vector<int> rp(A.nrows + 1), col(A.nnz);
    HIP_TRY(hipMemcpyAsync(rp.data(), A.row_ptr, sizeof(int) * (A.nrows + 1), hipMemcpyDeviceToHost, st));
    if (A.nnz) HIP_TRY(hipMemcpyAsync(col.data(), A.col, sizeof(int) * A.nnz, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int hb = 0;
    for (int i = 0; i < A.nrows; i++)
        for (int p = rp[i]; p < rp[i + 1]; p++) hb = std::max(hb, std::abs(col[p] - i));
    return hb;
}

template <class FT>
void scatter(const DevMatrix &A, int dense, long ld, int hb, FT *out, hipStream_t st) {
    const long n = A.nrows;
    hipLaunchKernelGGL((k_zero<FT>), dim3(egrid(n * ld)), dim3(LB), 0, st, n * ld, out);
    LAUNCH_CHECK();
    if (A.fmt == DevFmt::CRS)
        hipLaunchKernelGGL((k_scatter_crs<FT>), dim3(egrid(n)), dim3(LB), 0, st, n, dense, ld, hb, A.row_ptr, A.col,
                           (int)A.kind, (const void *)A.val, out);
    else
        hipLaunchKernelGGL((k_scatter_dia<FT>), dim3(egrid(n)), dim3(LB), 0, st, n, (long)A.ncols, ld, hb, A.ndiag,
                           A.offsets, (int)A.kind, (const void *)A.val, out);
    LAUNCH_CHECK();
}

template <class FT>
FT to_ft(std::complex<double> z) {
    if constexpr (Tr<FT>::is_complex) return FT(z.real(), z.imag());
    else return z.real();
}

template <class FT>
void build_typed(LuFactor &F, const DevMatrix &K, const DevMatrix *M, const DevMatrix *C,
                 const std::complex<double> *shift, hipStream_t st) {
    const long n = F.n, cnt = n * F.ld;
    const int dense = F.which == 2;
    FT *a = (FT *)F.lu;
    scatter<FT>(K, dense, F.ld, F.hb, a, st);
    if (shift) {
        const FT s = to_ft<FT>(*shift), s2 = s * s;
        FT *tc = nullptr, *tm = nullptr;
        struct Free {
            FT *&p;
            ~Free() { (void)hipFree(p); }
        } f1{tc}, f2{tm};
        if (C) {
            HIP_TRY(hipMalloc(&tc, sizeof(FT) * cnt));
            scatter<FT>(*C, dense, F.ld, F.hb, tc, st);
        }
        if (M) {
            HIP_TRY(hipMalloc(&tm, sizeof(FT) * cnt));
            scatter<FT>(*M, dense, F.ld, F.hb, tm, st);
        }
        if (M || C) {
            const int mode = (M && C) ? 0 : (C ? 1 : 2);
            hipLaunchKernelGGL((k_combine<FT>), dim3(egrid(cnt)), dim3(LB), 0, st, cnt, mode, s, s2, a, tc, tm);
            LAUNCH_CHECK();
        }
        if (C && !M) {          // mass matrix is identity: + s^2 on the diagonal
            hipLaunchKernelGGL((k_add_diag<FT>), dim3(egrid(n)), dim3(LB), 0, st, n, dense, F.ld, F.hb, s2, a);
            LAUNCH_CHECK();
        } else if (!M && !C) {  // - s on the diagonal
            hipLaunchKernelGGL((k_add_diag<FT>), dim3(egrid(n)), dim3(LB), 0, st, n, dense, F.ld, F.hb, FT(0) - s, a);
            LAUNCH_CHECK();
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    int *info = nullptr;
    HIP_TRY(hipMalloc(&info, sizeof(int)));
    struct FreeI {
        int *p;
        ~FreeI() { (void)hipFree(p); }
    } fi{info};
    hipLaunchKernelGGL((k_zero<int>), dim3(1), dim3(LB), 0, st, 1L, info);
    LAUNCH_CHECK();
    if (F.which == 3) {
        if (n > 1) hipLaunchKernelGGL((k_band_lu<FT>), dim3(1), dim3(FACT_THREADS), 0, st, n, F.hb, F.ld, a, info);
        LAUNCH_CHECK();
    } else {
        for (long k = 0; k < n; k++) {
            const long w = n - k - 1;
            if (w > 0) {
                const dim3 gd((unsigned)((w + LB - 1) / LB), (unsigned)std::min(w, 1024L));
                hipLaunchKernelGGL((k_dense_update<FT>), gd, dim3(LB), 0, st, n, k, a, info);
                LAUNCH_CHECK();
                hipLaunchKernelGGL((k_dense_lcol<FT>), dim3(egrid(w)), dim3(LB), 0, st, n, k, a, (const int *)info);
                LAUNCH_CHECK();
            } else {
                // the last pivot is still tested (RLU_decomp's loop runs to neq)
                hipLaunchKernelGGL((k_dense_update<FT>), dim3(1, 1), dim3(LB), 0, st, n, k, a, info);
                LAUNCH_CHECK();
            }
        }
    }
    int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (h) throw std::invalid_argument("Pivot equal to zero in LU decomposition");
}

}  // namespace

std::shared_ptr<LuFactor> lu_build(int which, const DevMatrix &K, const DevMatrix *M, const DevMatrix *C,
                                   const std::complex<double> *shift, bool cplx, hipStream_t st) {
    if (which != 2 && which != 3) throw std::invalid_argument("LU preconditioner: which_precon must be 2 or 3");
    if (K.nrows != K.ncols) throw std::invalid_argument("LU preconditioner: K must be square");
    for (const DevMatrix *X : {M, C})
        if (X && (X->nrows != K.nrows || X->ncols != K.ncols))
            throw std::invalid_argument("LU preconditioner: M and C must have K's shape");
    if (K.part) throw std::invalid_argument("LU preconditioner: not for batched systems");
    const bool any_complex = K.kind == ValKind::Complex || (M && M->kind == ValKind::Complex) ||
                             (C && C->kind == ValKind::Complex) || (shift && shift->imag() != 0.0);
    if (any_complex && !cplx) throw std::invalid_argument("REAL_PRECON: complex matrices or shift need COMPLEX_PRECON");
    if (which == 2)
        for (const DevMatrix *X : {&K, M, C})
            if (X && X->fmt != DevFmt::CRS)
                throw std::invalid_argument(
                    "dense LU preconditioner: CRS or COO matrices only (the reference's real_dense assembles no "
                    "entries of a CDS matrix and stops on the zero pivot)");
    auto F = std::make_shared<LuFactor>();
    F->which = which;
    F->n = K.nrows;
    F->cplx = cplx;
    if (which == 2) {
        if (F->n > 32768) throw std::invalid_argument("dense LU preconditioner: at most 32768 rows");
        F->hb = (int)(F->n - 1);
        F->ld = F->n;
    } else {
        int hb = band_of(K, st);
        if (C) hb = std::max(hb, band_of(*C, st));
        if (M) hb = std::max(hb, band_of(*M, st));
        F->hb = hb;
        F->ld = 2L * hb + 1;
    }
    if (!shift) M = C = nullptr;       // without a shift the reference factors K alone (the band still
                                       // spans M and C, user_module.f90:1423-1425)
    const size_t es = cplx ? sizeof(c128) : sizeof(double);
    HIP_TRY(hipMalloc(&F->lu, es * (size_t)F->n * F->ld));
    if (cplx) build_typed<c128>(*F, K, M, C, shift, st);
    else build_typed<double>(*F, K, M, C, shift, st);
    return F;
}

template <class T>
void lu_solve(const LuFactor &F, const T *v, T *w, hipStream_t st) {
    const int dense = F.which == 2;
    if (F.cplx) {
        if constexpr (Tr<T>::is_complex) {
            if (dense)
                hipLaunchKernelGGL((k_lu_solve<T, c128, 1024>), dim3(1), dim3(1024), 0, st, F.n, 1, F.hb, F.ld,
                                   (const c128 *)F.lu, v, w);
            else if (F.hb < WAVE)
                hipLaunchKernelGGL((k_band_solve_wave<T, c128>), dim3(1), dim3(WAVE), 0, st, F.n, F.hb, F.ld,
                                   (const c128 *)F.lu, v, w);
            else
                hipLaunchKernelGGL((k_lu_solve<T, c128, WAVE>), dim3(1), dim3(WAVE), 0, st, F.n, 0, F.hb, F.ld,
                                   (const c128 *)F.lu, v, w);
        } else {
            throw std::invalid_argument("complex LU preconditioner applied to a real vector");
        }
    } else {
        if (dense)
            hipLaunchKernelGGL((k_lu_solve<T, double, 1024>), dim3(1), dim3(1024), 0, st, F.n, 1, F.hb, F.ld,
                               (const double *)F.lu, v, w);
        else if (F.hb < WAVE)
            hipLaunchKernelGGL((k_band_solve_wave<T, double>), dim3(1), dim3(WAVE), 0, st, F.n, F.hb, F.ld,
                               (const double *)F.lu, v, w);
        else
            hipLaunchKernelGGL((k_lu_solve<T, double, WAVE>), dim3(1), dim3(WAVE), 0, st, F.n, 0, F.hb, F.ld,
                               (const double *)F.lu, v, w);
    }
    LAUNCH_CHECK();
}

template void lu_solve<double>(const LuFactor &, const double *, double *, hipStream_t);
template void lu_solve<c128>(const LuFactor &, const c128 *, c128 *, hipStream_t);

}  // namespace idrs
