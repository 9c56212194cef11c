// Sparse formats, SpMV kernels and the preconditioned operator.
//
// Reference behaviour restated here (paths relative to the reference's IDRS/):
//   CRS SpMV      MODULES/PP_MATRIX/user_module.f90:376-450 (rcrs_mv), :285-374 (ccrs_mv)
//   COO SpMV      user_module.f90:525-585 / :452-523
//   CDS SpMV      user_module.f90:676-827
//   diagonal      user_module.f90:960-1161, REAL_PRECON(1) :1327-1381
//   operator(*)   MODULES/PP_MATRIX/matrix_module.f90:135-264 (rmatvec, real_mv)
//   polpre        matrix_module.f90:266-304;  operator(/) :306-400
//   chebyshev / neumann / compute_alpha  matrix_module.f90:851-1057
//
// Design:
// - CRS rows are handled by groups of G lanes (G a power of two near the
//   average row length), so a wavefront streams the values and column
//   indices of 64/G consecutive rows as one coalesced stretch.
// - Group 1 uses a sliced-ELL copy instead, one lane per row, with each
//   slice of 64 rows stored entry-major.
// - DIA/CDS rows are one thread each, every diagonal a coalesced stream.
//   The stream bases are wave-uniform, and with diagonal scaling an all-ones
//   main diagonal is implicit.
// - Workgroups take contiguous slabs of rows and are remapped so that each
//   XCD streams one contiguous part of the matrix (L2 reuse of the x entries
//   neighbouring rows share).  One-thread-per-row products use up to 32768
//   short slabs.
// - The operator's follow-up work (Horner term, shadow-space dots, omega
//   dots, norms) is fused into the product as an epilogue.
#include "matrix.hpp"
#include "general.hpp"
#include "spmv_epi.hpp"
#include "ritz.hpp"
#include "lu.hpp"

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace idrs {

hipStream_t lib_stream() {
    static hipStream_t s = [] {
        hipStream_t t;
        HIP_TRY(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        return t;
    }();
    return s;
}

DevMatrix::~DevMatrix() {
    (void)hipFree(row_ptr);
    (void)hipFree(col);
    (void)hipFree(offsets);
    (void)hipFree(val);
    (void)hipFree(val32);
    (void)hipFree(diag);
    (void)hipFree(tperm);
}

Sell::~Sell() {
    (void)hipFree(sl_ptr);
    (void)hipFree(sl_col);
    (void)hipFree(sl_perm);
    (void)hipFree(sl_val);
    (void)hipFree(sl_val32);
}

Partition::~Partition() {
    (void)hipFree(row0);
    (void)hipFree(tile0);
    (void)hipFree(tile_sys);
    (void)hipFree(tile_r0);
}

Operator::~Operator() {
    (void)hipFree(pdiag);
    (void)hipFree(sval);
    (void)hipFree(sval32);
    (void)hipFree(sl_sval);
    (void)hipFree(sl_sval32);
    (void)hipFree(dV);
    (void)hipFree(dAV);
    (void)hipFree(dVtAV);
    (void)hipFree(gws);
}

SpmvTimer &spmv_timer() {
    static SpmvTimer t;
    return t;
}
SpmvTimer::~SpmvTimer() {}
double SpmvTimer::total_ms() const {
    double tot = 0.0;
    for (int i = 0; i + 1 < used; i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) tot += ms;
    }
    return tot;
}

// an event on the stream; the closing mark of an interval adds the launches
// it holds
static void timer_mark(hipStream_t st, int launches = 0) {
    SpmvTimer &t = spmv_timer();
    if (!t.enabled) return;
    t.counted += launches;
    if (t.used >= (int)t.ev.size()) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        t.ev.push_back(e);
    }
    HIP_TRY(hipEventRecord(t.ev[t.used++], st));
}

int vec_grid(long n) {
    long b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    if (b >= 64) b = (b / NUM_XCD) * NUM_XCD;
    return (int)b;
}

// blocks of a one-thread-per-row product (DIA, sliced CRS): one per 256 rows
// up to the row-slab cap, a multiple of the XCD count -- many short row slabs
// keep HBM busy to the end of the launch (epilogue partials: one per block)
static int wide_grid(long nrows, int cap) {
    long b = (nrows + SPMV_THREADS - 1) / SPMV_THREADS;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    if (b >= 64) b = (b / NUM_XCD) * NUM_XCD;
    return (int)b;
}

int spmv_grid(const DevMatrix &K) {
    if (K.fmt != DevFmt::CRS || K.group == 1) return wide_grid(K.nrows, K.row_slabs);
    long rows_per_unit = (K.fmt == DevFmt::CRS && K.group > 1) ? (long)(SPMV_THREADS / K.group) * 4 : SPMV_THREADS;
    long units = (K.nrows + rows_per_unit - 1) / rows_per_unit;
    long b = units;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    if (b >= 64) b = (b / NUM_XCD) * NUM_XCD;
    return (int)b;
}

template <class V>
static V *dev_upload(const V *h, size_t count) {
    V *d = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc(&d, sizeof(V) * count));
    if (h) HIP_TRY(hipMemcpy(d, h, sizeof(V) * count, hipMemcpyHostToDevice));
    return d;
}

// ---------------------------------------------------------------------------
// uploads
// ---------------------------------------------------------------------------
static void finish_crs(DevMatrix *M, const std::vector<int> &rp0, const std::vector<int> &col0,
                       const std::vector<double> &rv, const std::vector<c128> &cv) {
    long nnz = M->nnz;
    M->row_ptr = dev_upload(rp0.data(), rp0.size());
    M->col = dev_upload(col0.data(), (size_t)nnz);
    if (M->kind == ValKind::Complex) {
        M->val = dev_upload(cv.data(), (size_t)nnz);
    } else {
        M->val = dev_upload(rv.data(), (size_t)nnz);
    }
    // lanes per row: the smallest power of two >= the average row length
    double avg = M->nrows > 0 ? (double)nnz / M->nrows : 1.0;
    int g = 2;
    while (g < 64 && g < avg) g *= 2;
    M->group = g;
    // diagonal (real_diagonal / complex_diagonal), zeros -> 1
    if (M->kind == ValKind::Complex) {
        std::vector<c128> d(M->nrows, c128(0.0));
        for (int i = 0; i < M->nrows; i++)
            for (int j = rp0[i]; j < rp0[i + 1]; j++)
                if (col0[j] == i) d[i] = d[i] + cv[j];
        for (auto &v : d)
            if (v.re == 0.0 && v.im == 0.0) v = c128(1.0);
        M->diag = dev_upload(d.data(), d.size());
    } else {
        std::vector<double> d(M->nrows, 0.0);
        for (int i = 0; i < M->nrows; i++)
            for (int j = rp0[i]; j < rp0[i + 1]; j++)
                if (col0[j] == i) d[i] = d[i] + rv[j];
        for (auto &v : d)
            if (v == 0.0) v = 1.0;
        M->diag = dev_upload(d.data(), d.size());
    }
}

DevMatrix *make_crs(int nrows, int ncols, const int *row_ptr1, const int *col_ind1, const void *val,
                    ValKind kind, int orig_format) {
    if (nrows < 0 || ncols < 0) throw std::invalid_argument("negative dimension");
    if (row_ptr1[0] != 1) throw std::invalid_argument("row_ptr(1) must be 1 (1-based CRS)");
    long nnz = (long)row_ptr1[nrows] - 1;
    if (nnz < 0) throw std::invalid_argument("row_ptr(nrows+1) < 1");
    std::unique_ptr<DevMatrix> M(new DevMatrix);
    M->orig_format = orig_format;
    M->fmt = DevFmt::CRS;
    M->kind = kind == ValKind::Pattern ? ValKind::Real : kind;
    M->nrows = nrows;
    M->ncols = ncols;
    M->nnz = nnz;
    std::vector<int> rp0(nrows + 1), col0((size_t)nnz);
    for (int i = 0; i <= nrows; i++) {
        rp0[i] = row_ptr1[i] - 1;
        if (i > 0 && rp0[i] < rp0[i - 1]) throw std::invalid_argument("row_ptr not non-decreasing");
    }
    for (long j = 0; j < nnz; j++) {
        int c = col_ind1[j] - 1;
        if (c < 0 || c >= ncols) throw std::invalid_argument("col_ind out of range 1..ncols");
        col0[j] = c;
    }
    std::vector<double> rv;
    std::vector<c128> cv;
    if (M->kind == ValKind::Complex) {
        const c128 *v = (const c128 *)val;
        cv.assign(v, v + nnz);
    } else if (kind == ValKind::Pattern || !val) {
        rv.assign((size_t)nnz, 1.0);
    } else {
        const double *v = (const double *)val;
        rv.assign(v, v + nnz);
    }
    finish_crs(M.get(), rp0, col0, rv, cv);
    return M.release();
}

// COO -> CRS by a stable sort on the row index: within each row the entries
// keep their list order, so every row is summed in exactly the order of the
// reference's sequential scatter loop (user_module.f90:542-558).
DevMatrix *make_coo(int nrows, int ncols, long nnz, const int *indx1, const int *jndx1, const void *val,
                    ValKind kind) {
    if (nrows < 0 || ncols < 0 || nnz < 0) throw std::invalid_argument("negative dimension");
    std::unique_ptr<DevMatrix> M(new DevMatrix);
    M->orig_format = 3;
    M->fmt = DevFmt::CRS;
    M->kind = kind == ValKind::Pattern ? ValKind::Real : kind;
    M->nrows = nrows;
    M->ncols = ncols;
    M->nnz = nnz;
    std::vector<int> cnt(nrows + 1, 0);
    for (long k = 0; k < nnz; k++) {
        int i = indx1[k] - 1, j = jndx1[k] - 1;
        if (i < 0 || i >= nrows) throw std::invalid_argument("indx out of range 1..nrows");
        if (j < 0 || j >= ncols) throw std::invalid_argument("jndx out of range 1..ncols");
        cnt[i + 1]++;
    }
    std::vector<int> rp0(nrows + 1, 0);
    for (int i = 0; i < nrows; i++) rp0[i + 1] = rp0[i] + cnt[i + 1];
    std::vector<int> pos(rp0.begin(), rp0.end() - 1), col0((size_t)nnz);
    std::vector<double> rv;
    std::vector<c128> cv;
    if (M->kind == ValKind::Complex) cv.resize((size_t)nnz);
    else rv.resize((size_t)nnz);
    for (long k = 0; k < nnz; k++) {
        int i = indx1[k] - 1;
        int p = pos[i]++;
        col0[p] = jndx1[k] - 1;
        if (M->kind == ValKind::Complex) cv[p] = ((const c128 *)val)[k];
        else rv[p] = (kind == ValKind::Pattern || !val) ? 1.0 : ((const double *)val)[k];
    }
    finish_crs(M.get(), rp0, col0, rv, cv);
    return M.release();
}

static DevMatrix *make_dia_impl(int nrows, int ncols, int ndiag, const int *offsets, const void *vals,
                                ValKind kind, int orig_format, const int *diag_index_in_band) {
    if (ndiag < 1 || ndiag > MAX_DIAG) throw std::invalid_argument("ndiag out of range 1..1024");
    std::unique_ptr<DevMatrix> M(new DevMatrix);
    M->orig_format = orig_format;
    M->fmt = DevFmt::DIA;
    M->kind = kind == ValKind::Pattern ? ValKind::Real : kind;
    M->nrows = nrows;
    M->ncols = ncols;
    M->ndiag = ndiag;
    M->nnz = (long)ndiag * nrows;
    M->h_offsets.assign(offsets, offsets + ndiag);
    for (int d = 1; d < ndiag; d++)
        if (offsets[d] <= offsets[d - 1]) throw std::invalid_argument("DIA offsets must be strictly ascending");
    M->offsets = dev_upload(offsets, (size_t)ndiag);
    size_t cnt = (size_t)ndiag * nrows;
    int d0 = -1;
    for (int d = 0; d < ndiag; d++)
        if (offsets[d] == 0) d0 = d;
    (void)diag_index_in_band;
    if (M->kind == ValKind::Complex) {
        M->val = dev_upload((const c128 *)vals, cnt);
        std::vector<c128> dg(nrows, c128(1.0));
        if (d0 >= 0)
            for (int i = 0; i < nrows; i++) {
                c128 v = ((const c128 *)vals)[(size_t)d0 * nrows + i];
                dg[i] = (v.re == 0.0 && v.im == 0.0) ? c128(1.0) : v;
            }
        M->diag = dev_upload(dg.data(), dg.size());
    } else {
        std::vector<double> tmp;
        const double *src = (const double *)vals;
        if (kind == ValKind::Pattern || !vals) {
            tmp.assign(cnt, 1.0);
            src = tmp.data();
        }
        M->val = dev_upload(src, cnt);
        std::vector<double> dg(nrows, 1.0);
        if (d0 >= 0)
            for (int i = 0; i < nrows; i++) {
                double v = src[(size_t)d0 * nrows + i];
                dg[i] = v == 0.0 ? 1.0 : v;
            }
        M->diag = dev_upload(dg.data(), dg.size());
    }
    return M.release();
}

// cds_format: dense band (nrows, -hb:hb) column-major = DIA with all 2hb+1
// offsets in ascending order, the reference's diagonal loop order
// (user_module.f90:727-733).
DevMatrix *make_cds(int nrows, int ncols, int half_band, const void *band, ValKind kind) {
    if (half_band < 0 || 2 * half_band + 1 > MAX_DIAG) throw std::invalid_argument("half_band out of range");
    std::vector<int> off(2 * half_band + 1);
    for (int d = 0; d < 2 * half_band + 1; d++) off[d] = d - half_band;
    return make_dia_impl(nrows, ncols, 2 * half_band + 1, off.data(), band, kind, 4, nullptr);
}

DevMatrix *make_dia(int nrows, int ncols, int ndiag, const int *offsets, const void *vals, ValKind kind) {
    return make_dia_impl(nrows, ncols, ndiag, offsets, vals, kind, 5, nullptr);
}

// CRS: a group of G lanes per row (G >= the typical row length), so a
// wavefront covers 64/G consecutive rows whose entries are one contiguous,
// coalesced stretch of values and column indices.  The block's row pointers
// are staged in LDS once per chunk of rows, which leaves two dependent
// memory latencies per row (column index, then x).  Each group keeps U rows
// in flight: all their first-pass loads, and the P components of the P^T y
// epilogue, are issued before any is used; rows longer than G finish in a
// strided remainder loop.  The group's partial sums are combined
// by a butterfly inside the group, so every lane of the group holds y(row);
// for P^T y lane g takes the P components g, g+G, ... of its rows.
constexpr int CRS_CHUNK = 2048;   // rows whose row pointers sit in LDS at once

template <class VT, class T, int EPI, int NV, int G>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv_crs(SpmvArgs<VT, T> A, EpiArgs<T> E, const int *info) {
    if (info && *info >= 0) return;
    constexpr int GPB = SPMV_THREADS / G;
    constexpr int U = 4;
    // epilogue operands per row and lane: P components (PDOT), r (OMEGA)
    constexpr int PC = EPI == EPI_PDOT ? (NV + G - 1) / G : 1;
    __shared__ int s_rp[CRS_CHUNK + 1];
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    T e[EPI == EPI_PDOT ? PC : NV];
#pragma unroll
    for (int i = 0; i < (EPI == EPI_PDOT ? PC : NV); i++) e[i] = T(0);
    const int g = threadIdx.x % G, grp = threadIdx.x / G;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long n = A.nrows;
    const long r0 = n * lb / gridDim.x, r1 = n * (lb + 1) / gridDim.x;
    const T *__restrict__ x = A.x;
    const VT *__restrict__ val = A.val;
    const int *__restrict__ col = A.col;
    for (long c0 = r0; c0 < r1; c0 += CRS_CHUNK) {
        const long c1 = min(c0 + (long)CRS_CHUNK, r1);
        __syncthreads();
        for (long q = c0 + threadIdx.x; q <= c1; q += SPMV_THREADS) s_rp[q - c0] = A.rp[q];
        __syncthreads();
        for (long base = c0; base < c1; base += (long)GPB * U) {
            long row[U], rq[U];
            int a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                row[u] = base + grp + (long)u * GPB;
                const bool in = row[u] < c1;
                rq[u] = in ? row[u] : c0;
                a[u] = in ? s_rp[row[u] - c0] : 0;
                b[u] = in ? s_rp[row[u] - c0 + 1] : 0;
            }
            VT vv[U];
            int cc[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = a[u] + g;
                const bool ok = j < b[u];
                vv[u] = ok ? ldnt(&val[j]) : VT(0);
                cc[u] = ok ? ldnt(&col[j]) : -1;
            }
            // P components of the rows (P^T y) are issued with the matrix
            // loads; z and r are loaded at the finish, where the register
            // budget of the U rows in flight is lower
            T pe[U][PC];
            if constexpr (EPI == EPI_PDOT) {
#pragma unroll
                for (int u = 0; u < U; u++) epi_load<T, EPI, PC>(pe[u], E, rq[u], g, G);
            }
            T acc[U];
#pragma unroll
            for (int u = 0; u < U; u++) acc[u] = cc[u] >= 0 ? mulv(vv[u], x[cc[u]]) : T(0);
#pragma unroll
            for (int u = 0; u < U; u++)
                for (int j = a[u] + g + G; j < b[u]; j += G) acc[u] = fmav(ldnt(&val[j]), x[ldnt(&col[j])], acc[u]);
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int m = G / 2; m >= 1; m >>= 1) acc[u] = acc[u] + shfl_xor(acc[u], m);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (row[u] < c1) {
                    const T zq = A.z ? A.z[row[u]] : T(0);
                    const T yv = finish_value(A, row[u], acc[u], zq);
                    if (g == 0) A.y[row[u]] = yv;
                    if constexpr (EPI == EPI_PDOT) {
                        epi_accum<T, EPI, PC, PC>(e, pe[u], E, yv, g, G);
                    } else if (g == 0) {
                        epi_load<T, EPI, PC>(pe[u], E, row[u], 0, 1);
                        epi_accum<T, EPI, NV, PC>(e, pe[u], E, yv, 0, 1);
                    }
                }
        }
    }
    if constexpr (EPI == EPI_PDOT) epi_store_group<T, NV, G, PC>(e, E, s_red);
    else epi_store<T, EPI, NV>(e, E, s_red);
}

// DIA: one thread per row over a contiguous row range; every diagonal is a
// coalesced stream.  ND > 0 unrolls a known diagonal count.
template <class VT, class T, int EPI, int NV, int ND>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv_dia(SpmvArgs<VT, T> A, EpiArgs<T> E, const int *info) {
    if (info && *info >= 0) return;
    constexpr int PC = EPI == EPI_PDOT ? NV : 1;
    static_assert(ND <= 8, "known diagonal counts take their offsets by value");
    __shared__ int s_off[ND > 0 ? 1 : MAX_DIAG];
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    const int ndiag = ND > 0 ? ND : A.ndiag;
    if constexpr (ND == 0) {
        for (int d = threadIdx.x; d < ndiag; d += SPMV_THREADS) s_off[d] = A.offsets[d];
        __syncthreads();
    }
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long n = A.nrows, ncols = A.ncols;
    const long nchunk = (n + SPMV_THREADS - 1) / SPMV_THREADS;
    const long c0 = nchunk * lb / gridDim.x, c1 = nchunk * (lb + 1) / gridDim.x;
    const T *__restrict__ x = A.x;
    const VT *__restrict__ val = A.val;
    const int ud = A.ud;
    const T *zp = A.z;   // Horner vector; products without one read nothing here
    // diagonal d is stored at d - (d > ud); the implicit unit diagonal
    // (d == ud) enters as fmav(1, x, acc), the same operation as a stored
    // 1.0.  With a known diagonal count the offsets and value streams are
    // wave-uniform (scalar registers), rows index them in 32 bits.
    int offr[ND > 0 ? ND : 1];
    const VT *vb[ND > 0 ? ND : 1];
    if constexpr (ND > 0) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
            offr[d] = A.off8[d];
            vb[d] = val + (size_t)(d - (d > ud ? 1 : 0)) * (size_t)n;
        }
    }
    for (long c = c0; c < c1; c++) {
        const long row = c * SPMV_THREADS + threadIdx.x;
        if (row >= n) continue;
        T acc = T(0);
        T zq;
        T pe[PC];
        if constexpr (ND > 0) {
            // 32-bit byte offsets from the wave-uniform bases (one offset
            // register for every stream of the row instead of a 64-bit
            // address each); the launcher takes this path only while the
            // offsets fit: n * 8 * max(1, prow) < 2^32
            const unsigned bo = (unsigned)row * (unsigned)sizeof(T);
            const unsigned vbo = (unsigned)row * (unsigned)sizeof(VT);
            zq = zp ? *(const T *)((const char *)zp + bo) : T(0);
            if constexpr (EPI == EPI_PDOT) {
                const T *p = (const T *)((const char *)E.P + (unsigned)row * (unsigned)(E.prow * sizeof(T)));
                epi_load_row<T, PC>(pe, E, p);
            }
            VT vv[ND];
            T xx[ND];
#pragma unroll
            for (int d = 0; d < ND; d++) {
                // every load in bounds without a branch: the stored value
                // of row `row` exists on each diagonal, and a column outside
                // the matrix reads x(row) instead (the launcher takes this
                // path only for nrows <= ncols); both enter as zero
                const long cc = row + offr[d];
                const bool ok = cc >= 0 && cc < ncols;
                if constexpr (std::is_same<VT, T>::value) {
                    const unsigned xo = (unsigned)(ok ? cc : row) * (unsigned)sizeof(T);
                    const VT v = d == ud ? VT(1) : ldnt((const VT *)((const char *)vb[d] + vbo));
                    const T xv = *(const T *)((const char *)x + xo);
                    vv[d] = ok ? v : VT(0);
                    xx[d] = ok ? xv : T(0);
                } else {
                    // values of another type: loads under the column test, so
                    // that the conversion is not placed next to each load
                    // (which would wait for it before the next is issued)
                    vv[d] = !ok ? VT(0) : (d == ud ? VT(1) : ldnt((const VT *)((const char *)vb[d] + vbo)));
                    xx[d] = ok ? *(const T *)((const char *)x + (unsigned)cc * (unsigned)sizeof(T)) : T(0);
                }
            }
#pragma unroll
            for (int d = 0; d < ND; d++) acc = fmav(vv[d], xx[d], acc);
            const T yv = finish_value(A, row, acc, zq);
            *(T *)((char *)A.y + bo) = yv;
            if constexpr (EPI != EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
            epi_accum<T, EPI, NV, PC>(e, pe, E, yv, 0, 1);
            continue;
        } else {
            zq = zp ? zp[row] : T(0);
            if constexpr (EPI == EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
            for (int d = 0; d < ndiag; d++) {
                const long cc = row + s_off[d];
                const long sd = d - (d > ud ? 1 : 0);
                if (cc >= 0 && cc < ncols) acc = fmav(d == ud ? VT(1) : ldnt(&val[sd * n + row]), x[cc], acc);
            }
        }
        const T yv = finish_value(A, row, acc, zq);
        A.y[row] = yv;
        if constexpr (EPI != EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
        epi_accum<T, EPI, NV, PC>(e, pe, E, yv, 0, 1);
    }
    epi_store<T, EPI, NV>(e, E, s_red);
}

// Sliced copy of a CRS matrix (group 1): one thread per row.  Entry k of the
// row is at p0 + SELL_C k, so each load instruction of a wavefront reads
// SELL_C consecutive values (or column indices).  Entries go in chunks of 8:
// all value and index loads of a chunk, then all its x gathers; the row sum
// runs over k in CRS order (the reference's sequential row sum).
template <class VT, class T>
__device__ inline T sell_row(const SpmvArgs<VT, T> &A, long row, long p0, int w) {
    const VT *__restrict__ val = A.val;
    const int *__restrict__ scol = A.sl_col;
    const T *__restrict__ x = A.x;
    const int len = A.rp[row + 1] - A.rp[row];
    T acc = T(0);
    if constexpr (std::is_same<VT, float>::value) {
        // FP32 copy: one 8-byte load brings an entry's value and column
        // (4-byte loads of each would double the load instructions, which
        // bound this product); the same chunks and sum order as below
        if (A.pair32) {
            const unsigned long long *pr = (const unsigned long long *)val;
            for (int k0 = 0; k0 < w; k0 += 8) {
                float vv[8];
                int cc[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const bool ok = k0 + k < len;
                    const unsigned long long e = ok ? __builtin_nontemporal_load(&pr[p0 + (long)SELL_C * (k0 + k)]) : 0ull;
                    vv[k] = __uint_as_float((unsigned)(e & 0xffffffffull));
                    cc[k] = (int)(e >> 32);
                }
                T xx[8];
#pragma unroll
                for (int k = 0; k < 8; k++) xx[k] = k0 + k < len ? x[cc[k]] : T(0);
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if (k0 + k < len) acc = fmav(vv[k], xx[k], acc);
            }
            return acc;
        }
    }
    // chunks of 8 entries (w, the slice width, is wave-uniform): all value
    // and index loads of a chunk, then its x gathers, then the in-order sum
    for (int k0 = 0; k0 < w; k0 += 8) {
        VT vv[8];
        int cc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool ok = k0 + k < len;
            vv[k] = ok ? ldnt(&val[p0 + (long)SELL_C * (k0 + k)]) : VT(0);
            cc[k] = ok ? ldnt(&scol[p0 + (long)SELL_C * (k0 + k)]) : 0;
        }
        T xx[8];
#pragma unroll
        for (int k = 0; k < 8; k++) xx[k] = k0 + k < len ? x[cc[k]] : T(0);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (k0 + k < len) acc = fmav(vv[k], xx[k], acc);
    }
    return acc;
}

template <class VT, class T, int EPI, int NV>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv_sell(SpmvArgs<VT, T> A, EpiArgs<T> E, const int *info) {
    static_assert(SPMV_THREADS % SELL_C == 0 && SELL_C == WAVE, "one slice per wavefront");
    if (info && *info >= 0) return;
    constexpr int PC = EPI == EPI_PDOT ? NV : 1;
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long n = A.nrows;
    const long nchunk = (n + SPMV_THREADS - 1) / SPMV_THREADS;
    const long c0 = nchunk * lb / gridDim.x, c1 = nchunk * (lb + 1) / gridDim.x;
    const T *zp = A.z;   // Horner vector; products without one read nothing here
    const int wv = threadIdx.x / WAVE;
    for (long c = c0; c < c1; c++) {
        // the wavefront's slice (uniform: chunks start at multiples of SELL_C)
        const int slice = __builtin_amdgcn_readfirstlane((int)(c * (SPMV_THREADS / SELL_C) + wv));
        const long row = c * SPMV_THREADS + threadIdx.x;
        if (row >= n) continue;
        const long sp = A.sl_ptr[slice];
        const int w = (int)((A.sl_ptr[slice + 1] - sp) / SELL_C);
        const T zq = zp ? zp[row] : T(0);
        T pe[PC];
        if constexpr (EPI == EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
        const T acc = sell_row(A, row, sp + (row % SELL_C), w);
        const T yv = finish_value(A, row, acc, zq);
        A.y[row] = yv;
        if constexpr (EPI != EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
        epi_accum<T, EPI, NV, PC>(e, pe, E, yv, 0, 1);
    }
    epi_store<T, EPI, NV>(e, E, s_red);
}

template <class VT, class T, int EPI, int NV>
static void launch_spmv(const DevMatrix &K, const SpmvArgs<VT, T> &A, EpiArgs<T> E, const int *info,
                        hipStream_t st) {
    const int grid = spmv_grid(K);
    E.nparts = grid;
    dim3 gd(grid), bd(SPMV_THREADS);
    if (K.fmt == DevFmt::CRS && A.sl_ptr) {
        hipLaunchKernelGGL((k_spmv_sell<VT, T, EPI, NV>), gd, bd, 0, st, A, E, info);
    } else if (K.fmt == DevFmt::CRS) {
        // group 1 without sliced values (e.g. the FP32 copy): 2 lanes per row
        switch (K.group) {
        case 1:
        case 2: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 2>), gd, bd, 0, st, A, E, info); break;
        case 4: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 4>), gd, bd, 0, st, A, E, info); break;
        case 8: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 8>), gd, bd, 0, st, A, E, info); break;
        case 16: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 16>), gd, bd, 0, st, A, E, info); break;
        case 32: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 32>), gd, bd, 0, st, A, E, info); break;
        default: hipLaunchKernelGGL((k_spmv_crs<VT, T, EPI, NV, 64>), gd, bd, 0, st, A, E, info); break;
        }
    } else {
        // the unrolled kernels address rows by 32-bit byte offsets
        const unsigned long long rowb = sizeof(T) * (EPI == EPI_PDOT ? (unsigned long long)std::max(1, E.prow) : 1ull);
        const int nd =
            K.nrows <= K.ncols && (unsigned long long)K.ncols * rowb < (1ull << 32) ? K.ndiag : 0;
        switch (nd) {
        case 3: hipLaunchKernelGGL((k_spmv_dia<VT, T, EPI, NV, 3>), gd, bd, 0, st, A, E, info); break;
        case 5: hipLaunchKernelGGL((k_spmv_dia<VT, T, EPI, NV, 5>), gd, bd, 0, st, A, E, info); break;
        case 7: hipLaunchKernelGGL((k_spmv_dia<VT, T, EPI, NV, 7>), gd, bd, 0, st, A, E, info); break;
        default: hipLaunchKernelGGL((k_spmv_dia<VT, T, EPI, NV, 0>), gd, bd, 0, st, A, E, info); break;
        }
    }
    LAUNCH_CHECK();
}

template <class VT, class T>
static void dispatch_epi(const DevMatrix &K, const SpmvArgs<VT, T> &A, const EpiArgs<T> &E, int epi,
                         const int *info, hipStream_t st) {
    switch (epi) {
    case EPI_NONE: launch_spmv<VT, T, EPI_NONE, 1>(K, A, E, info, st); break;
    case EPI_NORM: launch_spmv<VT, T, EPI_NORM, 1>(K, A, E, info, st); break;
    case EPI_OMEGA: launch_spmv<VT, T, EPI_OMEGA, 3>(K, A, E, info, st); break;
    case EPI_PDOT:
        if (E.prow < 1 || E.prow > 16 || !E.P) throw std::invalid_argument("P^T y epilogue needs 1..16 rows of P");
        if (E.prow <= 1) launch_spmv<VT, T, EPI_PDOT, 1>(K, A, E, info, st);
        else if (E.prow <= 2) launch_spmv<VT, T, EPI_PDOT, 2>(K, A, E, info, st);
        else if (E.prow <= 4) launch_spmv<VT, T, EPI_PDOT, 4>(K, A, E, info, st);
        else if (E.prow <= 8) launch_spmv<VT, T, EPI_PDOT, 8>(K, A, E, info, st);
        else launch_spmv<VT, T, EPI_PDOT, 16>(K, A, E, info, st);
        break;
    default: throw std::invalid_argument("bad epilogue");
    }
}

// ---------------------------------------------------------------------------
// block products y_j = xs (K' x_j - seed x_j / sdiag) + a_j z_j over the
// columns of a block (the multishift solvers' per-shift products and
// scale-backs).  A row's entries are loaded once for all columns of the
// launch; each column's row sum and finish are those of k_spmv_dia /
// k_spmv_sell, so column j equals the one-column product of x_j bit for bit.
// ---------------------------------------------------------------------------
template <class VT, class T, int ND>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv_dia_cols(SpmvArgs<VT, T> A, ColArgs<T> C) {
    static_assert(ND > 0 && ND <= 8, "known diagonal counts only");
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long n = A.nrows, ncols = A.ncols;
    const long nchunk = (n + SPMV_THREADS - 1) / SPMV_THREADS;
    const long c0 = nchunk * lb / gridDim.x, c1 = nchunk * (lb + 1) / gridDim.x;
    const int ud = A.ud;
    for (long c = c0; c < c1; c++) {
        const long row = c * SPMV_THREADS + threadIdx.x;
        if (row >= n) continue;
        VT vv[ND];
        bool ok[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) {
            const long cc = row + A.off8[d];
            ok[d] = cc >= 0 && cc < ncols;
            const long sd = d - (d > ud ? 1 : 0);
            vv[d] = !ok[d] ? VT(0) : (d == ud ? VT(1) : ldnt(A.val + sd * n + row));
        }
#pragma unroll
        for (int j = 0; j < SPMV_COLS; j++) {
            if (j < C.nc) {
                const T *xj = C.x + j * C.ld;
                T xx[ND];
#pragma unroll
                for (int d = 0; d < ND; d++) xx[d] = ok[d] ? xj[row + A.off8[d]] : T(0);
                T acc = T(0);
#pragma unroll
                for (int d = 0; d < ND; d++) acc = fmav(vv[d], xx[d], acc);
                const T *zj = C.z ? C.z + j * C.ld : nullptr;
                const T zq = zj ? zj[row] : T(0);
                C.y[j * C.ld + row] = finish_col(A, xj, row, acc, zj, zq, C.a[j]);
            }
        }
    }
}

template <class VT, class T>
__global__ __launch_bounds__(SPMV_THREADS) void k_spmv_sell_cols(SpmvArgs<VT, T> A, ColArgs<T> C) {
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const long n = A.nrows;
    const long nchunk = (n + SPMV_THREADS - 1) / SPMV_THREADS;
    const long c0 = nchunk * lb / gridDim.x, c1 = nchunk * (lb + 1) / gridDim.x;
    const int wv = threadIdx.x / WAVE;
    const VT *__restrict__ val = A.val;
    const int *__restrict__ scol = A.sl_col;
    for (long c = c0; c < c1; c++) {
        const int slice = __builtin_amdgcn_readfirstlane((int)(c * (SPMV_THREADS / SELL_C) + wv));
        const long row = c * SPMV_THREADS + threadIdx.x;
        if (row >= n) continue;
        const long sp = A.sl_ptr[slice];
        const int w = (int)((A.sl_ptr[slice + 1] - sp) / SELL_C);
        const long p0 = sp + (row % SELL_C);
        const int len = A.rp[row + 1] - A.rp[row];
        T acc[SPMV_COLS];
#pragma unroll
        for (int j = 0; j < SPMV_COLS; j++) acc[j] = T(0);
        // the chunks of sell_row, each used for every column
        for (int k0 = 0; k0 < w; k0 += 8) {
            VT vv[8];
            int cc[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const bool ok = k0 + k < len;
                vv[k] = ok ? ldnt(&val[p0 + (long)SELL_C * (k0 + k)]) : VT(0);
                cc[k] = ok ? ldnt(&scol[p0 + (long)SELL_C * (k0 + k)]) : 0;
            }
#pragma unroll
            for (int j = 0; j < SPMV_COLS; j++) {
                if (j < C.nc) {
                    const T *xj = C.x + j * C.ld;
                    T xx[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) xx[k] = k0 + k < len ? xj[cc[k]] : T(0);
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        if (k0 + k < len) acc[j] = fmav(vv[k], xx[k], acc[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < SPMV_COLS; j++) {
            if (j < C.nc) {
                const T *zj = C.z ? C.z + j * C.ld : nullptr;
                const T zq = zj ? zj[row] : T(0);
                C.y[j * C.ld + row] = finish_col(A, C.x + j * C.ld, row, acc[j], zj, zq, C.a[j]);
            }
        }
    }
}

// one block product of at most SPMV_COLS columns; false when K's layout has
// no block kernel (generic DIA, CRS with lanes per row)
template <class VT, class T>
static bool launch_cols(const DevMatrix &K, const SpmvArgs<VT, T> &A, const ColArgs<T> &C, hipStream_t st) {
    dim3 gd(spmv_grid(K)), bd(SPMV_THREADS);
    if (K.fmt == DevFmt::CRS) {
        if (!A.sl_ptr || A.pair32) return false;
        hipLaunchKernelGGL((k_spmv_sell_cols<VT, T>), gd, bd, 0, st, A, C);
    } else {
        switch (K.ndiag) {
        case 3: hipLaunchKernelGGL((k_spmv_dia_cols<VT, T, 3>), gd, bd, 0, st, A, C); break;
        case 5: hipLaunchKernelGGL((k_spmv_dia_cols<VT, T, 5>), gd, bd, 0, st, A, C); break;
        case 7: hipLaunchKernelGGL((k_spmv_dia_cols<VT, T, 7>), gd, bd, 0, st, A, C); break;
        default: return false;
        }
    }
    LAUNCH_CHECK();
    return true;
}

// ---------------------------------------------------------------------------
// batched (tile) SpMV: one workgroup per tile of the block-diagonal matrix's
// partition.  Same per-row arithmetic and group layout as k_spmv_crs; the
// epilogue partials are per tile, and a tile whose system has finished
// returns at once.
// ---------------------------------------------------------------------------
struct TileArgs {
    const int *tile_sys;
    const long *tile_r0, *tile_r1;
    const int *info;
    int stride;
};

template <class VT, class T, int EPI, int NV, int G>
__global__ __launch_bounds__(SPMV_THREADS) void k_bspmv_crs(SpmvArgs<VT, T> A, EpiArgs<T> E, TileArgs ta) {
    const int tile = blockIdx.x;
    if (ta.info && ta.info[(long)ta.tile_sys[tile] * ta.stride] >= 0) return;
    constexpr int GPB = SPMV_THREADS / G;
    constexpr int U = 4;
    constexpr int PC = EPI == EPI_PDOT ? (NV + G - 1) / G : 1;
    __shared__ int s_rp[BATCH_TILE + 1];
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    T e[EPI == EPI_PDOT ? PC : NV];
#pragma unroll
    for (int i = 0; i < (EPI == EPI_PDOT ? PC : NV); i++) e[i] = T(0);
    const int g = threadIdx.x % G, grp = threadIdx.x / G;
    const long c0 = ta.tile_r0[tile], c1 = ta.tile_r1[tile];
    const T *__restrict__ x = A.x;
    const VT *__restrict__ val = A.val;
    const int *__restrict__ col = A.col;
    for (long q = c0 + threadIdx.x; q <= c1; q += SPMV_THREADS) s_rp[q - c0] = A.rp[q];
    __syncthreads();
    for (long base = c0; base < c1; base += (long)GPB * U) {
        long row[U], rq[U];
        int a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            row[u] = base + grp + (long)u * GPB;
            const bool in = row[u] < c1;
            rq[u] = in ? row[u] : c0;
            a[u] = in ? s_rp[row[u] - c0] : 0;
            b[u] = in ? s_rp[row[u] - c0 + 1] : 0;
        }
        VT vv[U];
        int cc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = a[u] + g;
            const bool ok = j < b[u];
            vv[u] = ok ? ldnt(&val[j]) : VT(0);
            cc[u] = ok ? ldnt(&col[j]) : -1;
        }
        T pe[U][PC];
        if constexpr (EPI == EPI_PDOT) {
#pragma unroll
            for (int u = 0; u < U; u++) epi_load<T, EPI, PC>(pe[u], E, rq[u], g, G);
        }
        T acc[U];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] = cc[u] >= 0 ? mulv(vv[u], x[cc[u]]) : T(0);
#pragma unroll
        for (int u = 0; u < U; u++)
            for (int j = a[u] + g + G; j < b[u]; j += G) acc[u] = fmav(ldnt(&val[j]), x[ldnt(&col[j])], acc[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int m = G / 2; m >= 1; m >>= 1) acc[u] = acc[u] + shfl_xor(acc[u], m);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (row[u] < c1) {
                const T zq = A.z ? A.z[row[u]] : T(0);
                const T yv = finish_value(A, row[u], acc[u], zq);
                if (g == 0) A.y[row[u]] = yv;
                if constexpr (EPI == EPI_PDOT) {
                    epi_accum<T, EPI, PC, PC>(e, pe[u], E, yv, g, G);
                } else if (g == 0) {
                    epi_load<T, EPI, PC>(pe[u], E, row[u], 0, 1);
                    epi_accum<T, EPI, NV, PC>(e, pe[u], E, yv, 0, 1);
                }
            }
    }
    if constexpr (EPI == EPI_PDOT) epi_store_group<T, NV, G, PC>(e, E, s_red);
    else epi_store<T, EPI, NV>(e, E, s_red);
}

// tile form of k_spmv_sell: a tile starts at any row; each wavefront walks
// whole slices and masks the rows outside the tile
template <class VT, class T, int EPI, int NV>
__global__ __launch_bounds__(SPMV_THREADS) void k_bspmv_sell(SpmvArgs<VT, T> A, EpiArgs<T> E, TileArgs ta) {
    const int tile = blockIdx.x;
    if (ta.info && ta.info[(long)ta.tile_sys[tile] * ta.stride] >= 0) return;
    constexpr int PC = EPI == EPI_PDOT ? NV : 1;
    __shared__ T s_red[NV * (SPMV_THREADS / WAVE)];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    const long c0 = ta.tile_r0[tile], c1 = ta.tile_r1[tile];
    const T *zp = A.z;
    // wavefront w takes the slices q0 + w, q0 + w + 4, ... that overlap the
    // tile (one slice per wavefront: uniform start and width); lanes whose
    // row lies outside the tile idle
    const int lane = threadIdx.x % WAVE;
    const long q0 = c0 / SELL_C, q1 = (c1 + SELL_C - 1) / SELL_C;
    for (long q = q0 + threadIdx.x / WAVE; q < q1; q += SPMV_THREADS / WAVE) {
        const int slice = __builtin_amdgcn_readfirstlane((int)q);
        const long row = (long)slice * SELL_C + lane;
        if (row < c0 || row >= c1) continue;
        const long sp = A.sl_ptr[slice];
        const int w = (int)((A.sl_ptr[slice + 1] - sp) / SELL_C);
        const T zq = zp ? zp[row] : T(0);
        T pe[PC];
        if constexpr (EPI == EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
        const T acc = sell_row(A, row, sp + (row % SELL_C), w);
        const T yv = finish_value(A, row, acc, zq);
        A.y[row] = yv;
        if constexpr (EPI != EPI_PDOT) epi_load<T, EPI, PC>(pe, E, row, 0, 1);
        epi_accum<T, EPI, NV, PC>(e, pe, E, yv, 0, 1);
    }
    epi_store<T, EPI, NV>(e, E, s_red);
}

template <class VT, class T, int EPI, int NV>
static void launch_bspmv(const DevMatrix &K, const SpmvArgs<VT, T> &A, EpiArgs<T> E, const TileCtx &tc,
                         hipStream_t st) {
    if (K.fmt != DevFmt::CRS) throw std::invalid_argument("batched systems must be stored as CRS");
    const Partition &P = *tc.part;
    TileArgs ta = tc.tile_sys ? TileArgs{tc.tile_sys, tc.tile_r0, tc.tile_r1, tc.info, tc.stride}
                              : TileArgs{P.tile_sys, P.tile_r0, P.tile_r0 + 1, tc.info, tc.stride};
    const int nt = tc.tile_sys ? tc.ntiles : P.ntiles;
    if (nt < 1) return;
    E.nparts = nt;
    dim3 gd(nt), bd(SPMV_THREADS);
    if (A.sl_ptr) {
        hipLaunchKernelGGL((k_bspmv_sell<VT, T, EPI, NV>), gd, bd, 0, st, A, E, ta);
        LAUNCH_CHECK();
        return;
    }
    switch (K.group) {
    case 1:
    case 2: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 2>), gd, bd, 0, st, A, E, ta); break;
    case 4: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 4>), gd, bd, 0, st, A, E, ta); break;
    case 8: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 8>), gd, bd, 0, st, A, E, ta); break;
    case 16: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 16>), gd, bd, 0, st, A, E, ta); break;
    case 32: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 32>), gd, bd, 0, st, A, E, ta); break;
    default: hipLaunchKernelGGL((k_bspmv_crs<VT, T, EPI, NV, 64>), gd, bd, 0, st, A, E, ta); break;
    }
    LAUNCH_CHECK();
}

template <class VT, class T>
static void dispatch_epi_tiles(const DevMatrix &K, const SpmvArgs<VT, T> &A, const EpiArgs<T> &E, int epi,
                               const TileCtx &tc, hipStream_t st) {
    if (!tc.part || K.part.get() != tc.part) throw std::invalid_argument("tile launch without the matrix's partition");
    switch (epi) {
    case EPI_NONE: launch_bspmv<VT, T, EPI_NONE, 1>(K, A, E, tc, st); break;
    case EPI_NORM: launch_bspmv<VT, T, EPI_NORM, 1>(K, A, E, tc, st); break;
    case EPI_OMEGA: launch_bspmv<VT, T, EPI_OMEGA, 3>(K, A, E, tc, st); break;
    case EPI_PDOT:
        if (E.prow < 1 || E.prow > 16 || !E.P) throw std::invalid_argument("P^T y epilogue needs 1..16 rows of P");
        if (E.prow <= 1) launch_bspmv<VT, T, EPI_PDOT, 1>(K, A, E, tc, st);
        else if (E.prow <= 2) launch_bspmv<VT, T, EPI_PDOT, 2>(K, A, E, tc, st);
        else if (E.prow <= 4) launch_bspmv<VT, T, EPI_PDOT, 4>(K, A, E, tc, st);
        else if (E.prow <= 8) launch_bspmv<VT, T, EPI_PDOT, 8>(K, A, E, tc, st);
        else launch_bspmv<VT, T, EPI_PDOT, 16>(K, A, E, tc, st);
        break;
    default: throw std::invalid_argument("bad epilogue");
    }
}

template <class T> static T to_T(std::complex<double> c);
template <> double to_T<double>(std::complex<double> c) { return c.real(); }
template <> float to_T<float>(std::complex<double> c) { return (float)c.real(); }
template <> c128 to_T<c128>(std::complex<double> c) { return c128(c.real(), c.imag()); }
template <> c64 to_T<c64>(std::complex<double> c) { return c64((float)c.real(), (float)c.imag()); }

template <class VT, class T>
static SpmvArgs<VT, T> base_args(const DevMatrix &K, const VT *val) {
    SpmvArgs<VT, T> A{};
    A.nrows = K.nrows;
    A.ncols = K.ncols;
    A.rp = K.row_ptr;
    A.col = K.col;
    A.val = val;
    A.ndiag = K.ndiag;
    A.offsets = K.offsets;
    for (int d = 0; d < 8; d++) A.off8[d] = d < (int)K.h_offsets.size() ? K.h_offsets[d] : 0;
    A.ud = INT_MAX;
    A.sl_ptr = nullptr;
    A.sl_col = nullptr;
    A.pair32 = 0;
    return A;
}

// group 1: point the product at the sliced copy of the values it reads
// (K's own or the operator's scaled ones); without one it stays on CRS
template <class VT, class T>
static void use_sliced(const Operator &op, SpmvArgs<VT, T> &A) {
    const DevMatrix &K = *op.K;
    if (K.fmt != DevFmt::CRS || K.group != 1 || !K.sell) return;
    constexpr bool lo = std::is_same<VT, float>::value;
    const void *v = op.sval ? (lo ? op.sl_sval32 : op.sl_sval) : (lo ? K.sell->sl_val32 : K.sell->sl_val);
    if (!v) return;
    A.val = (const VT *)v;
    A.sl_ptr = K.sell->sl_ptr;
    A.sl_col = K.sell->sl_col;
    A.pair32 = lo ? 1 : 0;
}

// The argument block of one product of op: its values (complex, FP32 copy
// or real; scaled or not), layout (sliced copy, implicit unit diagonal) and
// seed shift.  f receives it and fills the vector fields.
template <class T, class F>
static void with_args(const Operator &op, F &&f) {
    const DevMatrix &K = *op.K;
    const bool cv = op.values_complex();
    const bool has_seed = op.seed != std::complex<double>(0.0, 0.0);
    using R = typename Tr<T>::real;
    constexpr bool single = std::is_same<R, float>::value;
    if (has_seed && single) throw std::invalid_argument("seed shift in single precision is not supported");
    if (has_seed && !Tr<T>::is_complex && op.seed.imag() != 0.0)
        throw std::invalid_argument("complex seed on a real system");
    if (cv) {
        if constexpr (Tr<T>::is_complex && !single) {
            auto A = base_args<c128, T>(K, (const c128 *)op.values());
            if (op.sval && op.unit_diag >= 0) A.ud = op.unit_diag;
            use_sliced(op, A);
            A.has_seed = has_seed;
            A.sdiag = op.which_precon == 1 ? (const c128 *)op.pdiag : nullptr;
            A.seed = to_T<T>(op.seed);
            f(A);
        } else {
            throw std::invalid_argument("complex matrix values need complex128 vectors");
        }
    } else if (op.values32 && !single) {
        if constexpr (std::is_same<T, double>::value) {
            const float *v32 = (const float *)(op.sval ? op.sval32 : K.val32);
            if (!v32) throw std::invalid_argument("single-precision values not prepared");
            if (has_seed) throw std::invalid_argument("seed shift with single-precision values is not supported");
            auto A = base_args<float, T>(K, v32);
            if (op.sval && op.unit_diag >= 0) A.ud = op.unit_diag;
            use_sliced(op, A);
            A.has_seed = 0;
            A.sdiag = nullptr;
            A.seed = T(0);
            f(A);
        } else {
            throw std::invalid_argument("single-precision values with double vectors: real systems only");
        }
    } else {
        const R *vals = single ? (const R *)(op.sval ? op.sval32 : K.val32) : (const R *)op.values();
        if (!vals) throw std::invalid_argument("single-precision values not prepared");
        auto A = base_args<R, T>(K, vals);
        if (op.sval && op.unit_diag >= 0) A.ud = op.unit_diag;
        use_sliced(op, A);
        A.has_seed = has_seed;
        A.sdiag = op.which_precon == 1 ? (const R *)op.pdiag : nullptr;
        A.seed = to_T<T>(op.seed);
        f(A);
    }
}

template <class T>
void op_spmv(const Operator &op, const T *x, T xs, const T *z, T a, T *y, int epi, const EpiArgs<T> &E,
             const int *info, hipStream_t st, const TileCtx *tc) {
    const DevMatrix &K = *op.K;
    with_args<T>(op, [&](auto &A) {
        A.x = x; A.xs = xs; A.z = z; A.a = a; A.y = y;
        if (tc) dispatch_epi_tiles(K, A, E, epi, *tc, st);
        else dispatch_epi(K, A, E, epi, info, st);
    });
}

template <class T>
void op_spmv_cols(const Operator &op, int nc, long ld, const T *x, T xs, const T *z, const T *a, T *y,
                  hipStream_t st) {
    const DevMatrix &K = *op.K;
    EpiArgs<T> none{};
    for (int j0 = 0; j0 < nc; j0 += SPMV_COLS) {
        const int m = std::min(SPMV_COLS, nc - j0);
        bool done = false;
        with_args<T>(op, [&](auto &A) {
            A.xs = xs;
            ColArgs<T> C{};
            C.nc = m;
            C.ld = ld;
            C.x = x + (long)j0 * ld;
            C.z = z ? z + (long)j0 * ld : nullptr;
            C.y = y + (long)j0 * ld;
            for (int j = 0; j < m; j++) C.a[j] = a ? a[j0 + j] : T(0);
            done = launch_cols(K, A, C, st);
        });
        if (!done)
            for (int j = j0; j < j0 + m; j++)
                op_spmv<T>(op, x + (long)j * ld, xs, z ? z + (long)j * ld : nullptr, a ? a[j] : T(0),
                           y + (long)j * ld, EPI_NONE, none, nullptr, st);
    }
}

// polynomial coefficient as the vector type
template <class T>
static T coef(const Operator &op, int i) {
    return to_T<T>(op.alpha[i]);
}

// double and complex<double> vectors: the general path's types
template <class T>
constexpr bool gen_type() {
    return std::is_same<T, double>::value || std::is_same<T, c128>::value;
}

template <class T>
void op_apply(const Operator &op, const T *v, T *w, T *w1, T *w2, int epi, const EpiArgs<T> &E,
              const int *info, hipStream_t st, const TileCtx *tc) {
    if (op.general()) {
        if (tc) throw std::invalid_argument("batched solves take standard operators only");
        if constexpr (gen_type<T>()) gen_apply<T>(op, v, w, epi, E, info, st);
        else throw std::invalid_argument("mass, damping, block and deflated operators run in double precision");
        return;
    }
    const int d = op.degree;
    EpiArgs<T> none{};
    timer_mark(st);
    if (d == 0) {
        op_spmv<T>(op, v, T(1), nullptr, T(0), w, epi, E, info, st, tc);
        timer_mark(st, 1);
        return;
    }
    // Horner: w = alpha_d v; for i = d-1..0: w = K'(w) + alpha_i v; then K'(w)
    T *cur = w1, *nxt = w2;
    op_spmv<T>(op, v, coef<T>(op, d), v, coef<T>(op, d - 1), cur, EPI_NONE, none, info, st, tc);
    for (int i = d - 2; i >= 0; i--) {
        op_spmv<T>(op, cur, T(1), v, coef<T>(op, i), nxt, EPI_NONE, none, info, st, tc);
        std::swap(cur, nxt);
    }
    op_spmv<T>(op, cur, T(1), nullptr, T(0), w, epi, E, info, st, tc);
    timer_mark(st, d + 1);
}

template <class T>
void op_apply_cols(const Operator &op, int nc, long ld, const T *V, T *W, T *W1, T *W2, hipStream_t st) {
    if (op.general()) {
        EpiArgs<T> none{};
        for (int j = 0; j < nc; j++)
            op_apply<T>(op, V + (long)j * ld, W + (long)j * ld, W1, W2, EPI_NONE, none, nullptr, st);
        return;
    }
    const int d = op.degree;
    if (d == 0) {
        op_spmv_cols<T>(op, nc, ld, V, T(1), nullptr, nullptr, W, st);
        return;
    }
    std::vector<T> a((size_t)nc);
    auto all = [&](T c) {
        std::fill(a.begin(), a.end(), c);
        return a.data();
    };
    T *cur = W1, *nxt = W2;
    op_spmv_cols<T>(op, nc, ld, V, coef<T>(op, d), V, all(coef<T>(op, d - 1)), cur, st);
    for (int i = d - 2; i >= 0; i--) {
        op_spmv_cols<T>(op, nc, ld, cur, T(1), V, all(coef<T>(op, i)), nxt, st);
        std::swap(cur, nxt);
    }
    op_spmv_cols<T>(op, nc, ld, cur, T(1), nullptr, nullptr, W, st);
}

// ---------------------------------------------------------------------------
// block-diagonal batches
// ---------------------------------------------------------------------------
// rows [0, n) of one system's CRS arrays into the batch's: row pointers and
// column indices shifted by the system's first entry and first row
__global__ void k_shift_rows(long n, const int *__restrict__ rp, int ent0, int *__restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = rp[i] + ent0;
}

DevMatrix *make_blockdiag(int nsys, const DevMatrix *const *Ks) {
    if (nsys < 1) throw std::invalid_argument("a batch needs at least one system");
    const ValKind kind = Ks[0]->kind;
    long nrows = 0, nnz = 0;
    for (int b = 0; b < nsys; b++) {
        const DevMatrix *K = Ks[b];
        if (!K) throw std::invalid_argument("null system matrix");
        if (K->fmt != DevFmt::CRS) throw std::invalid_argument("batched systems must be CRS or COO matrices");
        if (K->nrows != K->ncols || K->nrows < 1) throw std::invalid_argument("batched systems must be square");
        if (K->kind != kind) throw std::invalid_argument("batched systems must all be real or all complex");
        nrows += K->nrows;
        nnz += K->nnz;
    }
    if (nrows > INT32_MAX || nnz > INT32_MAX) throw std::invalid_argument("batch too large for int32 indices");
    std::unique_ptr<DevMatrix> M(new DevMatrix);
    M->orig_format = 2;
    M->fmt = DevFmt::CRS;
    M->kind = kind;
    M->nrows = M->ncols = (int)nrows;
    M->nnz = nnz;
    const size_t vsz = kind == ValKind::Complex ? sizeof(c128) : sizeof(double);
    HIP_TRY(hipMalloc(&M->row_ptr, sizeof(int) * (nrows + 1)));
    HIP_TRY(hipMalloc(&M->col, sizeof(int) * (nnz ? nnz : 1)));
    HIP_TRY(hipMalloc(&M->val, vsz * (nnz ? nnz : 1)));
    HIP_TRY(hipMalloc(&M->diag, vsz * nrows));
    std::unique_ptr<Partition> P(new Partition);
    P->nsys = nsys;
    P->h_row0.assign(nsys + 1, 0);
    P->h_tile0.assign(nsys + 1, 0);
    std::vector<int> tile_sys;
    std::vector<long> tile_r0;
    long r0 = 0, e0 = 0;
    for (int b = 0; b < nsys; b++) {
        const DevMatrix *K = Ks[b];
        const long n = K->nrows, nz = K->nnz;
        const int g = vec_grid(n);
        // row pointers shifted by e0 (the final entry comes with the last system)
        hipLaunchKernelGGL(k_shift_rows, dim3(g), dim3(256), 0, 0, n + (b == nsys - 1 ? 1 : 0), K->row_ptr, (int)e0,
                           M->row_ptr + r0);
        LAUNCH_CHECK();
        if (nz > 0) {
            hipLaunchKernelGGL(k_shift_rows, dim3(vec_grid(nz)), dim3(256), 0, 0, nz, K->col, (int)r0, M->col + e0);
            LAUNCH_CHECK();
            HIP_TRY(hipMemcpy((char *)M->val + vsz * e0, K->val, vsz * nz, hipMemcpyDeviceToDevice));
        }
        HIP_TRY(hipMemcpy((char *)M->diag + vsz * r0, K->diag, vsz * n, hipMemcpyDeviceToDevice));
        P->h_row0[b] = r0;
        P->h_tile0[b] = (int)tile_sys.size();
        for (long t = r0; t < r0 + n; t += BATCH_TILE) {
            tile_sys.push_back(b);
            tile_r0.push_back(t);
        }
        r0 += n;
        e0 += nz;
    }
    P->h_row0[nsys] = r0;
    P->h_tile0[nsys] = (int)tile_sys.size();
    tile_r0.push_back(r0);
    P->ntiles = (int)tile_sys.size();
    HIP_TRY(hipDeviceSynchronize());
    P->row0 = dev_upload(P->h_row0.data(), P->h_row0.size());
    P->tile0 = dev_upload(P->h_tile0.data(), P->h_tile0.size());
    P->tile_sys = dev_upload(tile_sys.data(), tile_sys.size());
    P->tile_r0 = dev_upload(tile_r0.data(), tile_r0.size());
    double avg = (double)nnz / (double)nrows;
    int g = 2;
    while (g < 64 && g < avg) g *= 2;
    M->group = g;
    M->part = std::move(P);
    return M.release();
}

// ---------------------------------------------------------------------------
// transposes, value updates and the pattern gradient (differentiable solves)
// ---------------------------------------------------------------------------
DevMatrix *make_transpose(const DevMatrix &K) {
    if (K.fmt != DevFmt::CRS) throw std::invalid_argument("transpose needs a CRS/COO matrix");
    if (K.part) throw std::invalid_argument("transpose of a batch matrix is not offered");
    const int n = K.nrows, m = K.ncols;
    const long nnz = K.nnz;
    std::vector<int> rp(n + 1), col((size_t)nnz);
    HIP_TRY(hipMemcpy(rp.data(), K.row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (nnz) HIP_TRY(hipMemcpy(col.data(), K.col, sizeof(int) * nnz, hipMemcpyDeviceToHost));
    // counting sort by column; rows visited in order keep each new row sorted
    std::vector<int> trp(m + 1, 0), tcol((size_t)nnz), perm((size_t)nnz);
    for (long k = 0; k < nnz; k++) trp[col[k] + 1]++;
    for (int j = 0; j < m; j++) trp[j + 1] += trp[j];
    std::vector<int> pos(trp.begin(), trp.end() - 1);
    for (int i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int p = pos[col[k]]++;
            tcol[p] = i;
            perm[p] = k;
        }
    std::unique_ptr<DevMatrix> T(new DevMatrix);
    T->orig_format = 2;
    T->fmt = DevFmt::CRS;
    T->kind = K.kind;
    T->nrows = m;
    T->ncols = n;
    T->nnz = nnz;
    T->row_ptr = dev_upload(trp.data(), trp.size());
    T->col = dev_upload(tcol.data(), (size_t)nnz);
    T->tperm = dev_upload(perm.data(), (size_t)nnz);
    const size_t vsz = K.kind == ValKind::Complex ? sizeof(c128) : sizeof(double);
    HIP_TRY(hipMalloc(&T->val, vsz * (nnz ? nnz : 1)));
    HIP_TRY(hipMalloc(&T->diag, vsz * (m ? m : 1)));
    double avg = m > 0 ? (double)nnz / m : 1.0;
    int g = 2;
    while (g < 64 && g < avg) g *= 2;
    T->group = g;
    transpose_refresh(K, *T, 0);
    HIP_TRY(hipDeviceSynchronize());
    return T.release();
}

template <class V>
__global__ void k_gather(long n, const int *__restrict__ perm, const V *__restrict__ src, V *__restrict__ dst) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// d[i] = sum of the stored entries (i, i), zero -> 1 (real_diagonal /
// complex_diagonal, user_module.f90:960-1161)
template <class V>
__global__ void k_crs_diag(int n, const int *__restrict__ rp, const int *__restrict__ col, const V *__restrict__ val,
                           V *__restrict__ d) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        V acc = V(0);
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (col[k] == i) acc = acc + val[k];
        if (Tr<V>::absv(acc) == 0.0) acc = V(1.0);
        d[i] = acc;
    }
}

static void refresh_diag(DevMatrix &K, hipStream_t st) {
    const int g = vec_grid(K.nrows);
    if (K.kind == ValKind::Complex)
        hipLaunchKernelGGL(k_crs_diag<c128>, dim3(g), dim3(256), 0, st, K.nrows, K.row_ptr, K.col,
                           (const c128 *)K.val, (c128 *)K.diag);
    else
        hipLaunchKernelGGL(k_crs_diag<double>, dim3(g), dim3(256), 0, st, K.nrows, K.row_ptr, K.col,
                           (const double *)K.val, (double *)K.diag);
    LAUNCH_CHECK();
}

static void *sliced_copy(const Sell &S, const void *src, bool cplx, hipStream_t st, void *dst = nullptr);

// the sliced copy's values after K.val changed, gathered into the existing
// buffer (a free would synchronize the device on every refresh); the FP32
// copy is dropped and rebuilt when next needed
static void refresh_sliced(DevMatrix &K, hipStream_t st) {
    if (!K.sell) return;
    if (K.sell->sl_val32) {
        (void)hipFree(K.sell->sl_val32);
        K.sell->sl_val32 = nullptr;
    }
    K.sell->sl_val = sliced_copy(*K.sell, K.val, K.kind == ValKind::Complex, st, K.sell->sl_val);
}

void transpose_refresh(const DevMatrix &K, DevMatrix &KT, hipStream_t st) {
    if (!KT.tperm || KT.nnz != K.nnz || KT.nrows != K.ncols || KT.kind != K.kind)
        throw std::invalid_argument("not a transpose of this matrix");
    if (K.nnz) {
        const int g = vec_grid(K.nnz);
        if (K.kind == ValKind::Complex)
            hipLaunchKernelGGL(k_gather<c128>, dim3(g), dim3(256), 0, st, K.nnz, KT.tperm, (const c128 *)K.val,
                               (c128 *)KT.val);
        else
            hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(256), 0, st, K.nnz, KT.tperm, (const double *)K.val,
                               (double *)KT.val);
        LAUNCH_CHECK();
    }
    refresh_sliced(KT, st);
    if (KT.nrows) refresh_diag(KT, st);
}

void set_values(DevMatrix &K, const void *vals, bool device, hipStream_t st) {
    if (K.fmt != DevFmt::CRS) throw std::invalid_argument("set_values needs a CRS/COO matrix");
    if (K.orig_format == 3) throw std::invalid_argument("COO matrices are stored row-sorted; set CRS values");
    if (K.tperm) throw std::invalid_argument("refresh a transpose from its source matrix instead");
    const size_t vsz = K.kind == ValKind::Complex ? sizeof(c128) : sizeof(double);
    if (K.nnz)
        HIP_TRY(hipMemcpyAsync(K.val, vals, vsz * K.nnz, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    if (K.val32) {
        (void)hipFree(K.val32);
        K.val32 = nullptr;
    }
    refresh_sliced(K, st);
    if (K.nrows) refresh_diag(K, st);
}

template <class T, class V>
__global__ void k_pattern_grad(int n, const int *__restrict__ rp, const int *__restrict__ col,
                               const T *__restrict__ lam, const T *__restrict__ x, T *__restrict__ grad) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const T li = lam[i];
        for (int k = rp[i]; k < rp[i + 1]; k++) grad[k] = -(li * x[col[k]]);
    }
}

template <class T>
void pattern_grad(const DevMatrix &K, const T *lambda, const T *x, T *grad, hipStream_t st) {
    if (K.fmt != DevFmt::CRS) throw std::invalid_argument("pattern gradient needs a CRS/COO matrix");
    if (K.orig_format == 3) throw std::invalid_argument("pattern gradient is in CRS entry order; use a CRS matrix");
    if (!K.nrows) return;
    hipLaunchKernelGGL((k_pattern_grad<T, T>), dim3(vec_grid(K.nrows)), dim3(256), 0, st, K.nrows, K.row_ptr, K.col,
                       lambda, x, grad);
    LAUNCH_CHECK();
}
template void pattern_grad<double>(const DevMatrix &, const double *, const double *, double *, hipStream_t);
template void pattern_grad<c128>(const DevMatrix &, const c128 *, const c128 *, c128 *, hipStream_t);

// ---------------------------------------------------------------------------
// elementwise kernels: w = v / D  (rprecon), w = c*v
// ---------------------------------------------------------------------------
template <class T, class DT>
__global__ void k_divide(long n, const T *__restrict__ v, const DT *__restrict__ d, T *__restrict__ w) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        w[i] = divv(v[i], d[i]);
}
template <class T>
__global__ void k_scale_copy(long n, T c, const T *__restrict__ v, T *__restrict__ w) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        w[i] = c * v[i];
}

template <class T>
void op_precon(const Operator &op, const T *v, T *w, hipStream_t st) {
    if (op.general()) {
        if constexpr (gen_type<T>()) gen_precon<T>(op, v, w, st);
        else throw std::invalid_argument("mass, damping, block and deflated operators run in double precision");
        return;
    }
    long n = op.K->nrows;
    int g = vec_grid(n);
    if (op.which_precon == 1) {
        if constexpr (Tr<T>::is_complex) {
            using R = typename Tr<T>::real;
            if (op.pdiag_kind == ValKind::Complex) {
                if constexpr (std::is_same<R, double>::value)
                    hipLaunchKernelGGL((k_divide<T, c128>), dim3(g), dim3(256), 0, st, n, v, (const c128 *)op.pdiag, w);
                else
                    throw std::invalid_argument("complex diagonal in single precision");
            } else {
                if constexpr (std::is_same<R, double>::value)
                    hipLaunchKernelGGL((k_divide<T, double>), dim3(g), dim3(256), 0, st, n, v, (const double *)op.pdiag, w);
                else
                    throw std::invalid_argument("single-precision precon");
            }
        } else {
            if (op.pdiag_kind == ValKind::Complex) throw std::invalid_argument("complex diagonal on real vector");
            if constexpr (std::is_same<T, double>::value)
                hipLaunchKernelGGL((k_divide<double, double>), dim3(g), dim3(256), 0, st, n, v, (const double *)op.pdiag, w);
            else
                throw std::invalid_argument("single-precision precon");
        }
    } else {
        if (v != w) HIP_TRY(hipMemcpyAsync(w, v, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        return;
    }
    LAUNCH_CHECK();
}

template <class T>
void op_scaleback_with(const Operator &op, const std::vector<std::complex<double>> &alpha, const T *y, T *x, T *w1,
                       T *w2, hipStream_t st) {
    if (op.general()) {
        // without deflation (its step needs b and the shift: gen_ms_scaleback_col)
        if (op.ndefl > 0) throw std::invalid_argument("the scale-back of a deflated operator needs b");
        if constexpr (gen_type<T>()) gen_ms_scaleback_col<T>(op, alpha, op.seed, nullptr, y, x, st);
        else throw std::invalid_argument("mass, damping, block and deflated operators run in double precision");
        return;
    }
    const int d = (int)alpha.size() - 1;
    long n = op.K->nrows;
    EpiArgs<T> none{};
    const T *p = y;
    if (d > 0) {
        // polpre: w = alpha_d y; for i = d-1..0: w = K'(w) + alpha_i y
        int g = vec_grid(n);
        T *cur = w1, *nxt = w2;
        hipLaunchKernelGGL((k_scale_copy<T>), dim3(g), dim3(256), 0, st, n, to_T<T>(alpha[d]), y, cur);
        LAUNCH_CHECK();
        for (int i = d - 1; i >= 0; i--) {
            op_spmv<T>(op, cur, T(1), y, to_T<T>(alpha[i]), nxt, EPI_NONE, none, nullptr, st);
            std::swap(cur, nxt);
        }
        p = cur;
    }
    op_precon<T>(op, p, x, st);
}

template <class T>
void op_scaleback_cols(const Operator &op, int nc, long ld, const std::vector<std::vector<std::complex<double>>> &alpha,
                       const T *Y, T *X, T *W1, T *W2, hipStream_t st) {
    if ((int)alpha.size() < nc) throw std::invalid_argument("one coefficient set per column");
    if (op.general()) {
        for (int j = 0; j < nc; j++) op_scaleback_with<T>(op, alpha[j], Y + (long)j * ld, X + (long)j * ld, W1, W2, st);
        return;
    }
    const int d = nc > 0 ? (int)alpha[0].size() - 1 : 0;
    for (int j = 0; j < nc; j++)
        if ((int)alpha[j].size() != d + 1) throw std::invalid_argument("coefficient sets of unequal degree");
    const long n = op.K->nrows;
    const T *P = Y;
    if (d > 0) {
        // polpre on every column: W = alpha_d y; for i = d-1..0: W = K'(W) + alpha_i y
        const int g = vec_grid(n);
        T *cur = W1, *nxt = W2;
        for (int j = 0; j < nc; j++) {
            hipLaunchKernelGGL((k_scale_copy<T>), dim3(g), dim3(256), 0, st, n, to_T<T>(alpha[j][d]), Y + (long)j * ld,
                               cur + (long)j * ld);
            LAUNCH_CHECK();
        }
        std::vector<T> a((size_t)nc);
        for (int i = d - 1; i >= 0; i--) {
            for (int j = 0; j < nc; j++) a[j] = to_T<T>(alpha[j][i]);
            op_spmv_cols<T>(op, nc, ld, cur, T(1), Y, a.data(), nxt, st);
            std::swap(cur, nxt);
        }
        P = cur;
    }
    for (int j = 0; j < nc; j++) op_precon<T>(op, P + (long)j * ld, X + (long)j * ld, st);
}

template <class T>
void op_scaleback(const Operator &op, const T *y, T *x, T *w1, T *w2, T *w3, hipStream_t st) {
    (void)w3;
    if (op.degree > 0)
        op_scaleback_with<T>(op, std::vector<std::complex<double>>(op.alpha.begin(), op.alpha.begin() + op.degree + 1),
                             y, x, w1, w2, st);
    else op_scaleback_with<T>(op, {std::complex<double>(1.0, 0.0)}, y, x, w1, w2, st);
}

template <class T>
void user_spmv(const DevMatrix &K, const T *v, T *w, hipStream_t st) {
    Operator tmp;
    tmp.K = const_cast<DevMatrix *>(&K);
    EpiArgs<T> none{};
    op_spmv<T>(tmp, v, T(1), nullptr, T(0), w, EPI_NONE, none, nullptr, st);
    tmp.K = nullptr;
}

// sliced copies: dst[p] = src[perm[p]] (padding: zero)
template <class V>
__global__ void k_sell_gather(long nent, const int *__restrict__ perm, const V *__restrict__ src, V *__restrict__ dst) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < nent; p += (long)gridDim.x * blockDim.x) {
        const int q = perm[p];
        dst[p] = q >= 0 ? src[q] : V(0);
    }
}

static void *sliced_copy(const Sell &S, const void *src, bool cplx, hipStream_t st, void *dst) {
    const size_t vsz = cplx ? sizeof(c128) : sizeof(double);
    if (!dst) HIP_TRY(hipMalloc(&dst, vsz * (S.nent ? S.nent : 1)));
    const int g = vec_grid(S.nent);
    if (cplx)
        hipLaunchKernelGGL((k_sell_gather<c128>), dim3(g), dim3(256), 0, st, S.nent, S.sl_perm, (const c128 *)src,
                           (c128 *)dst);
    else
        hipLaunchKernelGGL((k_sell_gather<double>), dim3(g), dim3(256), 0, st, S.nent, S.sl_perm,
                           (const double *)src, (double *)dst);
    LAUNCH_CHECK();
    return dst;
}

// slice widths and the entry permutation on the host (row pointers read
// back once), column indices and values gathered on the device
void build_sell(DevMatrix &K, hipStream_t st) {
    if (K.fmt != DevFmt::CRS) throw std::invalid_argument("sliced copy of a non-CRS matrix");
    const long n = K.nrows;
    std::vector<int> rp(n + 1);
    HIP_TRY(hipMemcpy(rp.data(), K.row_ptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    std::unique_ptr<Sell> S(new Sell);
    S->nslices = (int)((n + SELL_C - 1) / SELL_C);
    std::vector<long> sp(S->nslices + 1, 0);
    for (int q = 0; q < S->nslices; q++) {
        int w = 0;
        for (long r = (long)q * SELL_C; r < std::min(n, (long)(q + 1) * SELL_C); r++) w = std::max(w, rp[r + 1] - rp[r]);
        sp[q + 1] = sp[q] + (long)w * SELL_C;
    }
    S->nent = sp[S->nslices];
    std::vector<int> perm(S->nent, -1);
    for (long r = 0; r < n; r++) {
        const long p0 = sp[r / SELL_C] + r % SELL_C;
        for (int k = 0; k < rp[r + 1] - rp[r]; k++) perm[p0 + (long)SELL_C * k] = rp[r] + k;
    }
    S->sl_ptr = dev_upload(sp.data(), sp.size());
    S->sl_perm = dev_upload(perm.data(), perm.size());
    HIP_TRY(hipMalloc(&S->sl_col, sizeof(int) * (S->nent ? S->nent : 1)));
    const int g = vec_grid(S->nent);
    hipLaunchKernelGGL((k_sell_gather<int>), dim3(g), dim3(256), 0, st, S->nent, S->sl_perm, (const int *)K.col,
                       S->sl_col);
    LAUNCH_CHECK();
    S->sl_val = sliced_copy(*S, K.val, K.kind == ValKind::Complex, st);
    HIP_TRY(hipStreamSynchronize(st));
    K.sell = std::move(S);
}

bool op_prepared(const Operator &op) {
    const DevMatrix &K = *op.K;
    const bool sliced = K.fmt != DevFmt::CRS || !K.sell || !op.sval || op.sl_sval;
    return sliced && (!op.general() || (op.gws && op.gws_rows == op.rows()));
}

void op_prepare(const Operator &op, hipStream_t st) {
    if (op_prepared(op)) return;
    const DevMatrix &K = *op.K;
    if (!(K.fmt != DevFmt::CRS || !K.sell || !op.sval || op.sl_sval))
        op.sl_sval = sliced_copy(*K.sell, op.sval, op.values_complex(), st);
    gen_prepare(op);
    HIP_TRY(hipStreamSynchronize(st));
}

// Autotuner of the product layout: the CRS kernel's lanes per row (the
// candidates around the average-row-length choice, and the sliced copy),
// then, for one-thread-per-row products, the number of row slabs (reps
// launches each, HIP events on the library stream); keeps the fastest.
// Explicit, so results stay reproducible: a group size changes the order in
// which a row's partial sums are combined, a slab count the order of the
// epilogue's block sums, i.e. the rounding.
int autotune_group(DevMatrix &K, int reps, hipStream_t st, std::vector<std::pair<int, float>> *times) {
    if (K.nrows < 1) return K.group;
    const bool cplx = K.kind == ValKind::Complex;
    const size_t vsz = cplx ? sizeof(c128) : sizeof(double);
    void *x = nullptr, *y = nullptr;
    HIP_TRY(hipMalloc(&x, vsz * (K.ncols ? K.ncols : 1)));
    HIP_TRY(hipMalloc(&y, vsz * K.nrows));
    HIP_TRY(hipMemsetAsync(x, 0, vsz * (K.ncols ? K.ncols : 1), st));
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    const int gin = K.group, sin = K.row_slabs;
    // ms per product with the current layout (one warm-up product first)
    auto timed = [&]() {
        if (cplx) user_spmv<c128>(K, (const c128 *)x, (c128 *)y, st);
        else user_spmv<double>(K, (const double *)x, (double *)y, st);
        HIP_TRY(hipEventRecord(a, st));
        for (int r = 0; r < reps; r++) {
            if (cplx) user_spmv<c128>(K, (const c128 *)x, (c128 *)y, st);
            else user_spmv<double>(K, (const double *)x, (double *)y, st);
        }
        HIP_TRY(hipEventRecord(b, st));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        return ms / reps;
    };
    int best = gin;
    try {
        if (K.fmt == DevFmt::CRS) {
            const int g0 = K.group == 1 ? 2 : K.group;
            float best_ms = 0.f;
            // the sliced copy is a candidate when its padding stays within
            // 1/4 of the stored entries (built for the timing, kept if chosen)
            bool sell_ok = false;
            if (!K.sell && K.nnz > 0) build_sell(K, st);
            if (K.sell) sell_ok = K.sell->nent <= K.nnz + K.nnz / 4;
            for (int g : {1, g0 / 2, g0, g0 * 2, g0 * 4}) {
                if (g == 1 ? !sell_ok : (g < 2 || g > 64)) continue;
                K.group = g;
                const float ms = timed();
                if (times) times->push_back({g, ms});
                if (best_ms == 0.f || ms < best_ms) {
                    best_ms = ms;
                    best = g;
                }
            }
            K.group = best;
            if (best != 1) K.sell.reset();
        }
        // row slabs of a one-thread-per-row product
        if (K.fmt != DevFmt::CRS || K.group == 1) {
            int best_s = K.row_slabs;
            float best_ms = 0.f;
            for (int sl : {4096, 8192, 16384, 32768}) {
                K.row_slabs = sl;
                const float ms = timed();
                if (times) times->push_back({-sl, ms});
                if (best_ms == 0.f || ms < best_ms) {
                    best_ms = ms;
                    best_s = sl;
                }
            }
            K.row_slabs = best_s;
        }
    } catch (...) {
        K.group = gin;
        K.row_slabs = sin;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        (void)hipFree(x);
        (void)hipFree(y);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(x);
    (void)hipFree(y);
    return best;
}

// ---------------------------------------------------------------------------
// operator construction
// ---------------------------------------------------------------------------
template <class VT, class DT>
__global__ void k_scale_cols_crs(long nnz, const int *__restrict__ col, const VT *__restrict__ val,
                                 const DT *__restrict__ d, VT *__restrict__ out) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += (long)gridDim.x * blockDim.x)
        out[j] = divv(val[j], d[col[j]]);
}
template <class VT, class DT>
__global__ void k_scale_cols_dia(int nrows, int ncols, int ndiag, const int *__restrict__ off,
                                 const VT *__restrict__ val, const DT *__restrict__ d, VT *__restrict__ out) {
    long tot = (long)nrows * ndiag;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
        int dd = (int)(q / nrows);
        long row = q - (long)dd * nrows;
        long c = row + off[dd];
        out[q] = (c >= 0 && c < ncols) ? divv(val[q], d[c]) : val[q];
    }
}
template <class T>
__global__ void k_add_scalar(long n, T *__restrict__ d, T c) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[i] = d[i] + c;
}
template <class VT, class OT>
__global__ void k_convert(long n, const VT *__restrict__ in, OT *__restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (OT)in[i];
}
// (float value, column) pairs of a sliced copy: low word the value's bits,
// high word the column
__global__ void k_pair32(long n, const double *__restrict__ val, const int *__restrict__ col,
                         unsigned long long *__restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (unsigned long long)__float_as_uint((float)val[i]) | ((unsigned long long)(unsigned)col[i] << 32);
}
__global__ void k_real_to_cplx(long n, const double *__restrict__ in, c128 *__restrict__ out) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = c128(in[i], 0.0);
}

// per block: 1 if some v[i] is not exactly 1 (complex: (1, +0))
__device__ inline bool exactly_one(double v) { return v == 1.0; }
__device__ inline bool exactly_one(c128 v) { return v.re == 1.0 && __double_as_longlong(v.im) == 0; }
template <class VT>
__global__ void k_not_one(long n, const VT *__restrict__ v, int *__restrict__ bad) {
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    bool b = false;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b = b || !exactly_one(v[i]);
    if (b) s_bad = 1;
    __syncthreads();
    if (threadIdx.x == 0) bad[blockIdx.x] = s_bad;
}

// DIA with diagonal scaling: K D^-1 has the main diagonal a_ii / a_ii,
// exactly 1 wherever a_ii is a nonzero without a shift.  When all of it is
// 1, the scaled values drop that diagonal (one value stream less per
// product) and the kernel multiplies by an implicit 1.
static void drop_unit_diagonal(Operator &op, bool cvals, hipStream_t st) {
    DevMatrix &K = *op.K;
    const long n = K.nrows;
    int z0 = -1;
    for (int d = 0; d < K.ndiag; d++)
        if (K.h_offsets[d] == 0) z0 = d;
    if (z0 < 0 || n < 1 || K.ndiag < 2) return;
    const size_t vsz = cvals ? sizeof(c128) : sizeof(double);
    const int g = std::min(vec_grid(n), 1024);
    int *bad = nullptr;
    HIP_TRY(hipMalloc(&bad, sizeof(int) * g));
    std::vector<int> hb(g);
    try {
        if (cvals)
            hipLaunchKernelGGL((k_not_one<c128>), dim3(g), dim3(256), 0, st, n, (const c128 *)op.sval + (long)z0 * n, bad);
        else
            hipLaunchKernelGGL((k_not_one<double>), dim3(g), dim3(256), 0, st, n, (const double *)op.sval + (long)z0 * n,
                               bad);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(hb.data(), bad, sizeof(int) * g, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipFree(bad);
        throw;
    }
    (void)hipFree(bad);
    for (int v : hb)
        if (v) return;
    void *packed = nullptr;
    HIP_TRY(hipMalloc(&packed, vsz * n * (K.ndiag - 1)));
    char *src = (char *)op.sval, *dst = (char *)packed;
    if (z0 > 0) HIP_TRY(hipMemcpyAsync(dst, src, vsz * n * z0, hipMemcpyDeviceToDevice, st));
    if (z0 < K.ndiag - 1)
        HIP_TRY(hipMemcpyAsync(dst + vsz * n * z0, src + vsz * n * (z0 + 1), vsz * n * (K.ndiag - 1 - z0),
                               hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    (void)hipFree(op.sval);
    op.sval = packed;
    op.unit_diag = z0;
}

long Operator::stored_values() const {
    const long full = K->fmt == DevFmt::CRS ? K->nnz : (long)K->ndiag * K->nrows;
    return (sval && unit_diag >= 0) ? full - K->nrows : full;
}

// REAL_PRECON/COMPLEX_PRECON(1, K[, shift]) (user_module.f90:1327-1381):
// D = diag(K) [- shift] with zeros already replaced by 1 in diag(K).  The
// values of K D^-1 are precomputed so that real_mv's t = v/D is folded into
// the product.
static void op_scale_values(Operator &op, bool cdiag, hipStream_t st);

void op_set_diag_precon(Operator &op, int which, const std::complex<double> *shift) {
    hipStream_t st = lib_stream();
    DevMatrix &K = *op.K;
    (void)hipFree(op.pdiag);
    (void)hipFree(op.sval);
    (void)hipFree(op.sval32);
    (void)hipFree(op.sl_sval);
    (void)hipFree(op.sl_sval32);
    op.pdiag = op.sval = op.sval32 = nullptr;
    op.sl_sval = op.sl_sval32 = nullptr;
    op.unit_diag = -1;
    op.which_precon = which;
    if (which != 1) return;
    if (K.nrows != K.ncols) throw std::invalid_argument("diagonal scaling needs a square matrix");
    long n = K.nrows;
    int g = vec_grid(n);
    bool cdiag = (K.kind == ValKind::Complex) || (shift && shift->imag() != 0.0);
    op.pdiag_kind = cdiag ? ValKind::Complex : ValKind::Real;
    if (cdiag) {
        HIP_TRY(hipMalloc(&op.pdiag, sizeof(c128) * (n ? n : 1)));
        if (K.kind == ValKind::Complex)
            HIP_TRY(hipMemcpyAsync(op.pdiag, K.diag, sizeof(c128) * n, hipMemcpyDeviceToDevice, st));
        else {
            hipLaunchKernelGGL(k_real_to_cplx, dim3(g), dim3(256), 0, st, n, (const double *)K.diag, (c128 *)op.pdiag);
            LAUNCH_CHECK();
        }
        if (shift) {
            hipLaunchKernelGGL((k_add_scalar<c128>), dim3(g), dim3(256), 0, st, n, (c128 *)op.pdiag,
                               c128(-shift->real(), -shift->imag()));
            LAUNCH_CHECK();
        }
    } else {
        HIP_TRY(hipMalloc(&op.pdiag, sizeof(double) * (n ? n : 1)));
        HIP_TRY(hipMemcpyAsync(op.pdiag, K.diag, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        if (shift) {
            hipLaunchKernelGGL((k_add_scalar<double>), dim3(g), dim3(256), 0, st, n, (double *)op.pdiag, -shift->real());
            LAUNCH_CHECK();
        }
    }
    op_scale_values(op, cdiag, st);
}

// K D^-1 (column-scaled values) for the fast path, from op.pdiag
static void op_scale_values(Operator &op, bool cdiag, hipStream_t st) {
    DevMatrix &K = *op.K;
    long cnt = K.fmt == DevFmt::CRS ? K.nnz : (long)K.ndiag * K.nrows;
    int gv = vec_grid(cnt);
    bool cvals = (K.kind == ValKind::Complex) || cdiag;
    HIP_TRY(hipMalloc(&op.sval, (cvals ? sizeof(c128) : sizeof(double)) * (cnt ? cnt : 1)));
    void *src = K.val;
    void *tmpc = nullptr;
    if (cvals && K.kind != ValKind::Complex) {
        HIP_TRY(hipMalloc(&tmpc, sizeof(c128) * (cnt ? cnt : 1)));
        hipLaunchKernelGGL(k_real_to_cplx, dim3(gv), dim3(256), 0, st, cnt, (const double *)K.val, (c128 *)tmpc);
        LAUNCH_CHECK();
        src = tmpc;
    }
    if (K.fmt == DevFmt::CRS) {
        if (cvals)
            hipLaunchKernelGGL((k_scale_cols_crs<c128, c128>), dim3(gv), dim3(256), 0, st, cnt, K.col,
                               (const c128 *)src, (const c128 *)op.pdiag, (c128 *)op.sval);
        else
            hipLaunchKernelGGL((k_scale_cols_crs<double, double>), dim3(gv), dim3(256), 0, st, cnt, K.col,
                               (const double *)src, (const double *)op.pdiag, (double *)op.sval);
    } else {
        if (cvals)
            hipLaunchKernelGGL((k_scale_cols_dia<c128, c128>), dim3(gv), dim3(256), 0, st, K.nrows, K.ncols,
                               K.ndiag, K.offsets, (const c128 *)src, (const c128 *)op.pdiag, (c128 *)op.sval);
        else
            hipLaunchKernelGGL((k_scale_cols_dia<double, double>), dim3(gv), dim3(256), 0, st, K.nrows, K.ncols,
                               K.ndiag, K.offsets, (const double *)src, (const double *)op.pdiag, (double *)op.sval);
    }
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(st));
    (void)hipFree(tmpc);
    if (K.fmt != DevFmt::CRS) drop_unit_diagonal(op, cvals, st);
    else op_prepare(op, st);
}

// A%P given by its diagonal (REAL_PRECON / COMPLEX_PRECON(1, K, M, C,
// shift) of other matrices, user_module.f90:1327-1381, 1465-1530)
void op_set_diag_values(Operator &op, const std::vector<std::complex<double>> &d, bool cdiag) {
    hipStream_t st = lib_stream();
    const long n = op.K->nrows;
    if ((long)d.size() != n) throw std::invalid_argument("preconditioner size does not match the system");
    (void)hipFree(op.pdiag);
    (void)hipFree(op.sval);
    (void)hipFree(op.sval32);
    (void)hipFree(op.sl_sval);
    (void)hipFree(op.sl_sval32);
    op.pdiag = op.sval = op.sval32 = nullptr;
    op.sl_sval = op.sl_sval32 = nullptr;
    op.unit_diag = -1;
    op.which_precon = 1;
    op.pdiag_kind = cdiag ? ValKind::Complex : ValKind::Real;
    if (cdiag) {
        std::vector<c128> h(n);
        for (long i = 0; i < n; i++) h[i] = c128(d[i].real(), d[i].imag());
        HIP_TRY(hipMalloc(&op.pdiag, sizeof(c128) * (n ? n : 1)));
        HIP_TRY(hipMemcpy(op.pdiag, h.data(), sizeof(c128) * n, hipMemcpyHostToDevice));
    } else {
        std::vector<double> h(n);
        for (long i = 0; i < n; i++) h[i] = d[i].real();
        HIP_TRY(hipMalloc(&op.pdiag, sizeof(double) * (n ? n : 1)));
        HIP_TRY(hipMemcpy(op.pdiag, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    op_scale_values(op, cdiag, st);
}

void op_make_lowprec(Operator &op) {
    hipStream_t st = lib_stream();
    DevMatrix &K = *op.K;
    if (op.values_complex()) throw std::invalid_argument("single-precision copy of complex values");
    long cnt = op.sval ? op.stored_values() : (K.fmt == DevFmt::CRS ? K.nnz : (long)K.ndiag * K.nrows);
    int gv = vec_grid(cnt);
    if (op.sval) {
        if (!op.sval32) {
            HIP_TRY(hipMalloc(&op.sval32, sizeof(float) * (cnt ? cnt : 1)));
            hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(256), 0, st, cnt, (const double *)op.sval,
                               (float *)op.sval32);
            LAUNCH_CHECK();
        }
    } else if (!K.val32) {
        HIP_TRY(hipMalloc(&K.val32, sizeof(float) * (cnt ? cnt : 1)));
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(256), 0, st, cnt, (const double *)K.val,
                           (float *)K.val32);
        LAUNCH_CHECK();
    }
    // float copies of the sliced values (group 1)
    if (K.fmt == DevFmt::CRS && K.sell) {
        op_prepare(op, st);
        const long ne = K.sell->nent;
        const int gs = vec_grid(ne);
        void *&dst = op.sval ? op.sl_sval32 : K.sell->sl_val32;
        const void *src = op.sval ? op.sl_sval : K.sell->sl_val;
        if (!dst && src) {
            // (value, column) pairs, 8 bytes per entry (see sell_row)
            HIP_TRY(hipMalloc(&dst, sizeof(unsigned long long) * (ne ? ne : 1)));
            hipLaunchKernelGGL(k_pair32, dim3(gs), dim3(256), 0, st, ne, (const double *)src, K.sell->sl_col,
                               (unsigned long long *)dst);
            LAUNCH_CHECK();
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// polynomial coefficients (host), restating matrix_module.f90:851-1057 in
// the scalar type of the caller (real arithmetic for real_preconditioner).
// ---------------------------------------------------------------------------
template <class S>
static void compute_alpha_t(int degree, const S *H, int ldh, int ncol, S *alpha) {
    if (degree == 0) {
        alpha[0] = S(1.0);
        return;
    }
    if (ldh < degree + 2 || ncol < degree + 1)
        throw std::invalid_argument("Hessenberg matrix too small for polynomial degree");
    const int m = degree + 2;
    std::vector<S> r((size_t)m * m, S(0.0));
    auto R = [&](int i, int j) -> S & { return r[(size_t)i + (size_t)m * j]; };
    R(0, 0) = S(1.0);
    for (int i = 1; i <= degree + 1; i++) {
        for (int c = 1; c <= degree + 1; c++) R(i, c) = R(i - 1, c - 1);
        for (int j = 0; j <= i - 1; j++) {
            S h = H[j + (size_t)ldh * (i - 1)];
            for (int c = 0; c <= degree + 1; c++) R(i, c) = R(i, c) - h * R(j, c);
        }
        S hd = H[i + (size_t)ldh * (i - 1)];
        for (int c = 0; c <= degree + 1; c++) R(i, c) = R(i, c) / hd;
    }
    for (int c = 0; c <= degree; c++) alpha[c] = -R(degree + 1, c + 1);
}

template <class S>
static void chebyshev_t(int degree, const S foci[2], S *alpha) {
    const int ld = degree + 2, nc = degree + 1;
    std::vector<S> J((size_t)ld * nc, S(0.0));
    auto JJ = [&](int i, int j) -> S & { return J[(size_t)(i - 1) + (size_t)ld * (j - 1)]; };
    S theta = (foci[0] + foci[1]) / S(2.0);
    S delta = (foci[0] - foci[1]) / S(2.0);
    S sigma = theta / delta;
    JJ(1, 1) = theta;
    JJ(2, 1) = -theta;
    S rho = S(1.0) / sigma;
    for (int i = 2; i <= degree + 1; i++) {
        JJ(i - 1, i) = -delta * rho / S(2.0);
        JJ(i, i) = delta * sigma;
        JJ(i + 1, i) = -delta * sigma + delta * rho / S(2.0);
        rho = S(1.0) / (S(2.0) * sigma - rho);
    }
    compute_alpha_t<S>(degree, J.data(), ld, nc, alpha);
}

template <class S>
static void neumann_t(int degree, S center, S *alpha) {
    const int ld = degree + 2, nc = degree + 1;
    std::vector<S> J((size_t)ld * nc, S(0.0));
    for (int i = 1; i <= degree + 1; i++) {
        J[(size_t)(i - 1) + (size_t)ld * (i - 1)] = center;
        J[(size_t)i + (size_t)ld * (i - 1)] = -center;
    }
    compute_alpha_t<S>(degree, J.data(), ld, nc, alpha);
}

void compute_alpha(int degree, const std::complex<double> *H, int ldh, int ncol, std::complex<double> *alpha) {
    compute_alpha_t<std::complex<double>>(degree, H, ldh, ncol, alpha);
}
void chebyshev(int degree, const std::complex<double> foci[2], std::complex<double> *alpha) {
    chebyshev_t<std::complex<double>>(degree, foci, alpha);
}
void neumann(int degree, std::complex<double> center, std::complex<double> *alpha) {
    neumann_t<std::complex<double>>(degree, center, alpha);
}

// coefficient computation behind real_preconditioner / complex_preconditioner
// (matrix_module.f90:551-722, 851-1057), in real arithmetic unless complex_poly.
// Without foci (Chebyshev) or center (Neumann) but with H, they come from
// analyse_spectrum(H) (matrix_module.f90:600-619, 687-706), whose Ritz values
// are those of the leading ncol x ncol block.
void poly_coeffs(int kind, int degree, const std::complex<double> *foci, const std::complex<double> *center,
                 const std::complex<double> *H, int ldh, int ncol, bool complex_poly, std::complex<double> *out) {
    const bool from_h = H && ((kind == POLY_NEUMANN && !center) || (kind == POLY_CHEBYSHEV && !foci));
    if (from_h && ldh < ncol) throw std::invalid_argument("H has fewer rows than columns");
    if (!complex_poly) {
        std::vector<double> a((size_t)degree + 1), Hr;
        if (from_h) {
            Hr.resize((size_t)ldh * ncol);
            for (size_t q = 0; q < Hr.size(); q++) Hr[q] = H[q].real();
        }
        if (kind == POLY_NEUMANN) {
            double c = 1.0;
            if (center) c = center->real();
            else if (from_h) ritz::analyse_spectrum(ncol, Hr.data(), ldh, nullptr, &c, nullptr);
            neumann_t<double>(degree, c, a.data());
        } else if (kind == POLY_CHEBYSHEV) {
            double f[2] = {0.25, 1.75};
            if (foci) {
                f[0] = foci[0].real();
                f[1] = foci[1].real();
            } else if (from_h) {
                ritz::analyse_spectrum(ncol, Hr.data(), ldh, f, nullptr, nullptr);
            }
            chebyshev_t<double>(degree, f, a.data());
        } else if (kind == POLY_HESSENBERG) {
            if (!H) throw std::invalid_argument("Hessenberg preconditioner needs H");
            std::vector<double> Hr((size_t)ldh * ncol);
            for (size_t q = 0; q < Hr.size(); q++) Hr[q] = H[q].real();
            compute_alpha_t<double>(degree, Hr.data(), ldh, ncol, a.data());
        } else {
            throw std::invalid_argument("unknown polynomial kind");
        }
        for (int i = 0; i <= degree; i++) out[i] = a[i];
    } else {
        if (kind == POLY_NEUMANN) {
            std::complex<double> c = 1.0;
            if (center) c = *center;
            else if (from_h) ritz::analyse_spectrum(ncol, H, ldh, nullptr, &c, nullptr);
            neumann_t<std::complex<double>>(degree, c, out);
        } else if (kind == POLY_CHEBYSHEV) {
            std::complex<double> f[2] = {0.25, 1.75};
            if (foci) {
                f[0] = foci[0];
                f[1] = foci[1];
            } else if (from_h) {
                ritz::analyse_spectrum(ncol, H, ldh, f, nullptr, nullptr);
            }
            chebyshev_t<std::complex<double>>(degree, f, out);
        } else if (kind == POLY_HESSENBERG) {
            if (!H) throw std::invalid_argument("Hessenberg preconditioner needs H");
            compute_alpha_t<std::complex<double>>(degree, H, ldh, ncol, out);
        } else {
            throw std::invalid_argument("unknown polynomial kind");
        }
    }
}

// real_preconditioner / complex_preconditioner (matrix_module.f90:551-722):
// foci and center are shifted by the seed before the coefficients are formed.
void op_set_poly(Operator &op, int kind, int degree, const std::complex<double> *foci,
                 const std::complex<double> *center, const std::complex<double> *H, int ldh, int ncol,
                 bool complex_poly) {
    if (degree < 0) throw std::invalid_argument("negative degree");
    op.poly_kind = kind;
    op.complex_poly = complex_poly;
    if (kind == POLY_NONE) {
        op.degree = 0;
        op.alpha.assign(1, std::complex<double>(1.0));
        return;
    }
    const std::complex<double> sd = complex_poly ? op.seed : std::complex<double>(op.seed.real(), 0.0);
    std::complex<double> f[2], c;
    if (foci) {
        f[0] = foci[0] - sd;
        f[1] = foci[1] - sd;
    }
    if (center) c = *center - sd;
    std::vector<std::complex<double>> a((size_t)degree + 1);
    poly_coeffs(kind, degree, foci ? f : nullptr, center ? &c : nullptr, H, ldh, ncol, complex_poly, a.data());
    op.degree = degree;
    op.alpha = a;
}

// explicit instantiations
#define INST(T)                                                                                              \
    template void op_spmv<T>(const Operator &, const T *, T, const T *, T, T *, int, const EpiArgs<T> &,     \
                             const int *, hipStream_t, const TileCtx *);                                     \
    template void op_apply<T>(const Operator &, const T *, T *, T *, T *, int, const EpiArgs<T> &,           \
                              const int *, hipStream_t, const TileCtx *);                                    \
    template void op_precon<T>(const Operator &, const T *, T *, hipStream_t);                               \
    template void op_scaleback<T>(const Operator &, const T *, T *, T *, T *, T *, hipStream_t);             \
    template void op_scaleback_with<T>(const Operator &, const std::vector<std::complex<double>> &, const T *, T *,  \
                                       T *, T *, hipStream_t);                                                 \
    template void op_spmv_cols<T>(const Operator &, int, long, const T *, T, const T *, const T *, T *, hipStream_t); \
    template void op_apply_cols<T>(const Operator &, int, long, const T *, T *, T *, T *, hipStream_t);         \
    template void op_scaleback_cols<T>(const Operator &, int, long,                                             \
                                       const std::vector<std::vector<std::complex<double>>> &, const T *, T *,  \
                                       T *, T *, hipStream_t);                                                  \
    template void user_spmv<T>(const DevMatrix &, const T *, T *, hipStream_t);
INST(double)
INST(c128)
INST(float)
INST(c64)
#undef INST

}  // namespace idrs
