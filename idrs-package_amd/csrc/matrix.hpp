// Device-resident user matrices (CRS / COO / CDS / DIA) and the solver
// operator A = K D^-1 p(K D^-1) (diagonal scaling + polynomial
// preconditioner), the objects behind the C-ABI handles.
#pragma once

#include "common.hpp"

#include <complex>
#include <memory>
#include <vector>

namespace idrs {

enum class ValKind { Pattern = 0, Real = 1, Complex = 2 };

// Storage after upload: CRS (CRS and COO inputs) or DIA (CDS and DIA inputs).
enum class DevFmt { CRS = 0, DIA = 1 };

constexpr int SPMV_THREADS = 256;
constexpr int MAX_DIAG = 1024;       // offsets held in LDS by the DIA kernel

// Row partition of a block-diagonal matrix into independent systems, cut
// into tiles (row ranges inside one system) for the batched kernels: one
// workgroup per tile, so every block-level partial sum belongs to exactly
// one system.
constexpr int BATCH_TILE = 1024;   // rows per tile
struct Partition {
    int nsys = 0, ntiles = 0;
    std::vector<long> h_row0;      // nsys+1: first row of each system
    std::vector<int> h_tile0;      // nsys+1: first tile of each system
    long *row0 = nullptr;          // device copies of the above
    int *tile0 = nullptr;
    int *tile_sys = nullptr;       // ntiles: system of each tile
    long *tile_r0 = nullptr;       // ntiles+1: first row of each tile (tiles are contiguous)
    ~Partition();
};

// Tile launch of the batched solver: per-tile partials, and per-system
// early exit once a system has finished (info: the State.info of system 0,
// stride: ints between consecutive systems' info; nullptr: never skip).
struct TileCtx {
    const Partition *part = nullptr;
    const int *info = nullptr;
    int stride = 0;
    // optional subset of the partition's tiles (e.g. those of unfinished
    // systems): system, first row and end row of each of ntiles tiles
    int ntiles = 0;
    const int *tile_sys = nullptr;
    const long *tile_r0 = nullptr, *tile_r1 = nullptr;
};

// Sliced ELL copy of a CRS matrix, used when group == 1 (one lane per row):
// slices of SELL_C consecutive rows, each stored entry-major at the slice's
// width (its longest row), so entry k of row r sits at
// sl_ptr[r / SELL_C] + SELL_C k + r % SELL_C and every load of a wavefront
// reads SELL_C consecutive entries.  Row lengths still come from row_ptr.
constexpr int SELL_C = 64;
struct Sell {
    int nslices = 0;
    long nent = 0;                 // sliced entries, padding included
    long *sl_ptr = nullptr;        // nslices+1
    int *sl_col = nullptr;         // nent (padding: 0, never read)
    int *sl_perm = nullptr;        // nent: CRS entry of each sliced entry, -1 for padding
    void *sl_val = nullptr;        // K's values (double or c128) in sliced order
    void *sl_val32 = nullptr;      // (float value, column) pairs of sl_val (op_make_lowprec)
    ~Sell();
};

struct DevMatrix {
    int orig_format = 0;     // user_module matrix_format: 2 CRS, 3 COO, 4 CDS, 5 DIA
    DevFmt fmt = DevFmt::CRS;
    ValKind kind = ValKind::Real;
    int nrows = 0, ncols = 0;
    long nnz = 0;            // stored entries (CRS) or ndiag*nrows (DIA)
    // CRS (0-based on device)
    int *row_ptr = nullptr, *col = nullptr;
    int group = 8;           // lanes that share one row in the CRS kernel; 1: the sliced copy
    std::unique_ptr<Sell> sell;   // built for group 1 (autotune_group)
    // most row slabs (workgroups) of a one-thread-per-row product (DIA, the
    // sliced copy): 4096, 8192, 16384 or 32768, chosen by autotune_group
    int row_slabs = 32768;
    // DIA
    int ndiag = 0;
    int *offsets = nullptr;
    std::vector<int> h_offsets;
    // values: double (Real/Pattern) or c128 (Complex); CRS: nnz, DIA: ndiag*nrows
    void *val = nullptr;
    // optional float copy (mixed precision), created on demand
    void *val32 = nullptr;
    // main diagonal (real_diagonal / complex_diagonal, zeros -> 1), device, as c128
    // or double depending on kind
    void *diag = nullptr;
    // set for block-diagonal batches (make_blockdiag)
    std::unique_ptr<Partition> part;
    // set for transposes (make_transpose): val[p] = source val[tperm[p]]
    int *tperm = nullptr;
    ~DevMatrix();
};

// Polynomial preconditioner kinds (matrix_module real_preconditioner)
enum { POLY_NONE = 0, POLY_NEUMANN = 1, POLY_CHEBYSHEV = 2, POLY_HESSENBERG = 3 };

struct LuFactor;

struct Operator {
    DevMatrix *K = nullptr;       // not owned
    int which_precon = 0;         // 0 identity, 1 diagonal scaling (A%P)
    // diagonal of the preconditioner (D = diag(K) [+ shift terms]), device
    void *pdiag = nullptr;        // double or c128 like K
    ValKind pdiag_kind = ValKind::Real;
    // values of K D^-1 (column scaled) when which_precon == 1, same layout as K
    void *sval = nullptr;
    void *sval32 = nullptr;       // float copy of the scaled values
    // DIA only: position (in K's diagonal list) of the main diagonal when its
    // scaled values are all exactly 1; sval then omits that diagonal and the
    // SpMV multiplies by an implicit 1 (-1: every diagonal stored)
    int unit_diag = -1;
    // sval in K's sliced order (K.sell), made by op_prepare
    mutable void *sl_sval = nullptr;
    mutable void *sl_sval32 = nullptr;   // (float value, column) pairs of sl_sval (op_make_lowprec)
    int degree = 0;
    int poly_kind = POLY_NONE;
    std::vector<std::complex<double>> alpha;   // alpha_p(0:degree)
    std::complex<double> seed{0.0, 0.0};       // A%rseed / A%cseed
    bool complex_poly = false;                 // coefficients set by complex_preconditioner
    // products of a double-vector solve read the FP32 copy of the (scaled)
    // values (set by the mixed-precision driver around its inner solves)
    bool values32 = false;
    // the generalized operator (matrix_module.f90:169-264, 306-400,
    // 1284-1623): any of these sends products through general.hip
    const DevMatrix *Mm = nullptr;   // A%M (mass_matrix), not owned
    const DevMatrix *Cm = nullptr;   // A%C (damping_matrix), not owned
    bool block = false;              // A%block_system: vectors of 2 K%nrows
    bool multishift = false;         // A%multishift: deflation projector I - V V^H
    int ndefl = 0;                   // A%n_deflation
    bool defl_complex = false;       // dV, dAV, dVtAV hold c128 (else double)
    void *dV = nullptr, *dAV = nullptr, *dVtAV = nullptr;   // rows x nd, rows x nd, nd x nd (device)
    // which_precon 2 / 3: the LU factors of REAL_PRECON(2|3, ...) (shared
    // between operators built from one preconditioner), and A%shift_invert
    // with A%rshift / A%cshift (user_module.f90:1383-1459,
    // matrix_module.f90:306-549, 724-806)
    std::shared_ptr<LuFactor> lu;
    bool shift_invert = false;
    std::complex<double> si_shift{0.0, 0.0};
    mutable void *gws = nullptr;     // workspace of the general path (gen_prepare)
    mutable long gws_rows = 0;
    bool general() const { return Mm || Cm || block || ndefl > 0 || lu || shift_invert; }
    long rows() const { return block ? 2L * K->nrows : K->nrows; }
    ~Operator();
    const void *values() const { return sval ? sval : K->val; }
    // number of value entries the SpMV reads (sval, sval32 or K's)
    long stored_values() const;
    bool values_complex() const {
        return (K->kind == ValKind::Complex) || (sval && pdiag_kind == ValKind::Complex);
    }
};

// epilogue kinds of the SpMV kernels
enum { EPI_NONE = 0, EPI_PDOT = 1, EPI_OMEGA = 2, EPI_NORM = 3 };

template <class T>
struct EpiArgs {
    const T *P = nullptr;       // shadow space PT(s,n): row q at P[q*s .. q*s+prow)
    int s = 0, prow = 0, conjP = 0;
    const T *r = nullptr;       // residual for EPI_OMEGA
    T *partials = nullptr;      // [v * nparts + block]
    int nparts = 0;
};

// Kernel timing (events around the dominant SpMV launches) for bench.
// HIP events around each operator application (its d+1 back-to-back SpMV
// launches, so one event pair's overhead is shared by all of them)
struct SpmvTimer {
    bool enabled = false;
    std::vector<hipEvent_t> ev;
    int used = 0;
    int counted = 0;   // SpMV launches inside the timed intervals
    void reset() { used = counted = 0; }
    double total_ms() const;
    int launches() const { return counted; }
    ~SpmvTimer();
};
SpmvTimer &spmv_timer();

// y = xs * (K' x) [+ a z] with epilogue; K' = scaled values when the operator
// has diagonal scaling.  T is the vector type.  'info' (device) makes the
// kernel a no-op once a solver has finished (nullptr: always run).
// With a TileCtx the product runs as one workgroup per tile of the matrix's
// partition (epilogue partials per tile; finished systems skipped).
template <class T>
void op_spmv(const Operator &op, const T *x, T xs, const T *z, T a, T *y, int epi, const EpiArgs<T> &E,
             const int *info, hipStream_t st, const TileCtx *tc = nullptr);

// w = A*v (rmatvec: Horner polynomial then system matvec), with the epilogue
// applied to the last SpMV.  w1, w2: scratch vectors of n.
template <class T>
void op_apply(const Operator &op, const T *v, T *w, T *w1, T *w2, int epi, const EpiArgs<T> &E,
              const int *info, hipStream_t st, const TileCtx *tc = nullptr);

// w = v / op (rprecon): divide by the preconditioner diagonal, or copy.
template <class T>
void op_precon(const Operator &op, const T *v, T *w, hipStream_t st);

// x = scaleback(op, y) = polpre(y) / D
template <class T>
void op_scaleback(const Operator &op, const T *y, T *x, T *w1, T *w2, T *w3, hipStream_t st);

// x = p(K') y / D with the given coefficients alpha(0:degree) in place of
// the operator's own (the per-shift scale-back of the multishift solvers)
template <class T>
void op_scaleback_with(const Operator &op, const std::vector<std::complex<double>> &alpha, const T *y, T *x, T *w1,
                       T *w2, hipStream_t st);

// Block forms over the nc columns of V, X, Y, W (column stride ld), one read
// of the matrix per up to 8 columns; column j is bit-identical to the
// one-column call on column j.  Layouts without a block kernel (generic DIA,
// CRS with lanes per row) run the one-column product per column.
//   op_spmv_cols:      y_j = xs * (K' x_j) + a[j] z_j  (z nullptr: no z term)
//   op_apply_cols:     W = A V (op_apply per column); W1, W2: blocks like V
//   op_scaleback_cols: X(:,j) = p_j(K') Y(:,j) / D, p_j's coefficients alpha[j]
constexpr int SPMV_COLS = 8;   // columns per block-product launch
template <class T>
void op_spmv_cols(const Operator &op, int nc, long ld, const T *x, T xs, const T *z, const T *a, T *y,
                  hipStream_t st);
template <class T>
void op_apply_cols(const Operator &op, int nc, long ld, const T *V, T *W, T *W1, T *W2, hipStream_t st);
template <class T>
void op_scaleback_cols(const Operator &op, int nc, long ld, const std::vector<std::vector<std::complex<double>>> &alpha,
                       const T *Y, T *X, T *W1, T *W2, hipStream_t st);

// plain user-matrix product w = K v (no scaling, no polynomial)
template <class T>
void user_spmv(const DevMatrix &K, const T *v, T *w, hipStream_t st);

// grid size for the streaming vector kernels
int vec_grid(long n);
int spmv_grid(const DevMatrix &K);

// uploads
DevMatrix *make_crs(int nrows, int ncols, const int *row_ptr1, const int *col_ind1, const void *val,
                    ValKind kind, int orig_format);
DevMatrix *make_coo(int nrows, int ncols, long nnz, const int *indx1, const int *jndx1, const void *val,
                    ValKind kind);
DevMatrix *make_cds(int nrows, int ncols, int half_band, const void *band, ValKind kind);
DevMatrix *make_dia(int nrows, int ncols, int ndiag, const int *offsets, const void *vals, ValKind kind);
// times the CRS product for lanes-per-row choices around K.group, and the
// sliced copy (group 1) when its padding is small, and keeps the fastest;
// for one-thread-per-row products (DIA, the sliced copy) then the row-slab
// count.  Returns the group; times: (group or -slabs, ms per product)
int autotune_group(DevMatrix &K, int reps, hipStream_t st, std::vector<std::pair<int, float>> *times);
// builds / drops K's sliced copy (group 1)
void build_sell(DevMatrix &K, hipStream_t st);
// makes the operator's sliced scaled values when K has a sliced copy; call
// before products are captured into a graph
void op_prepare(const Operator &op, hipStream_t st);
// true when op_prepare has nothing left to build for op
bool op_prepared(const Operator &op);
// block-diagonal CRS matrix of nsys square CRS systems (copies), with its
// Partition
DevMatrix *make_blockdiag(int nsys, const DevMatrix *const *Ks);
// CRS of K^T (rows of K^T list K's entries of that column in increasing row
// order), with the value permutation kept for transpose_refresh
DevMatrix *make_transpose(const DevMatrix &K);
// KT.val = K.val permuted (after K's values changed)
void transpose_refresh(const DevMatrix &K, DevMatrix &KT, hipStream_t st);
// replaces the stored values of a CRS matrix (same pattern) and recomputes
// its diagonal; vals is a device or host array of nnz values
void set_values(DevMatrix &K, const void *vals, bool device, hipStream_t st);
// grad[k] = -lambda[i] * x[j] for every stored entry k = (i, j) of K
template <class T>
void pattern_grad(const DevMatrix &K, const T *lambda, const T *x, T *grad, hipStream_t st);

// operator construction
void op_set_diag_precon(Operator &op, int which, const std::complex<double> *shift);
// A%P = a diagonal given on the host (which_precon 1)
void op_set_diag_values(Operator &op, const std::vector<std::complex<double>> &d, bool cdiag);
void op_set_poly(Operator &op, int kind, int degree, const std::complex<double> *foci,
                 const std::complex<double> *center, const std::complex<double> *H, int ldh, int ncol,
                 bool complex_poly);
void op_make_lowprec(Operator &op);
void poly_coeffs(int kind, int degree, const std::complex<double> *foci, const std::complex<double> *center,
                 const std::complex<double> *H, int ldh, int ncol, bool complex_poly, std::complex<double> *alpha);

// polynomial coefficient computation (host, restating matrix_module)
void compute_alpha(int degree, const std::complex<double> *H, int ldh, int ncol,
                   std::complex<double> *alpha);
void chebyshev(int degree, const std::complex<double> foci[2], std::complex<double> *alpha);
void neumann(int degree, std::complex<double> center, std::complex<double> *alpha);

hipStream_t lib_stream();

}  // namespace idrs
