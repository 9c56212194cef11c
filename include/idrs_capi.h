/*
 * idrs_capi.h - C ABI of the MI355X-native IDR(s) library (libidrs_hip.so).
 *
 * This is the boundary the Fortran drop-in modules (idrs-package_amd/fortran)
 * bind with iso_c_binding, and that Python binds with ctypes.  Each entry
 * point names the reference interface it replaces (paths relative to the
 * reference's IDRS/ directory).
 *
 * Conventions
 *   - Index arrays are 1-based, as the reference's converters take them.
 *   - Complex arrays are interleaved (re, im) doubles, i.e. Fortran
 *     complex(kind=real64) / C double _Complex.
 *   - 2-D arrays are Fortran column-major.
 *   - Vectors are HOST pointers unless idrs_opts_t.device_pointers is set
 *     (then b, x, x0, U0 and PT_raw are device pointers on the current HIP
 *     device, and the solve runs with inputs already resident in HBM).
 *   - Every function returns IDRS_OK or a negative status; idrs_last_error()
 *     gives the message.  The reference's "error stop" conditions (negative
 *     tolerance, short resvec, bad dimensions) return IDRS_ERR_ARG.
 *   - One host thread per device; calls are ordered on the library stream.
 */
#ifndef IDRS_CAPI_H
#define IDRS_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IDRS_OK 0
#define IDRS_ERR_ARG (-1)
#define IDRS_ERR_HIP (-2)
#define IDRS_ERR_UNSUPPORTED (-3)
#define IDRS_ERR_INTERNAL (-4)

typedef struct idrs_user_matrix_s *idrs_user_matrix_t; /* type(user_matrix) */
typedef struct idrs_matrix_s *idrs_matrix_t;           /* type(matrix)      */

/* Message of the last failed call on this thread (empty if none). */
const char *idrs_last_error(void);
/* Library version (major*10000 + minor*100 + patch). */
int idrs_version(void);

/* ---------------------------------------------------------------------
 * type(user_matrix) converters
 * ------------------------------------------------------------------- */

/* A = crs_format(nrows, ncols, row_ptr, col_ind, rval)
   replaces PP_MATRIX/user_module.f90:829-863 (crs_format); rval NULL gives
   the pattern matrix (all values 1). */
int idrs_crs_format_f64(int nrows, int ncols, const int *row_ptr, const int *col_ind, const double *rval,
                        idrs_user_matrix_t *A);
/* crs_format(..., cval=) - complex values, user_module.f90:829-863 */
int idrs_crs_format_c128(int nrows, int ncols, const int *row_ptr, const int *col_ind, const double *cval,
                         idrs_user_matrix_t *A);
/* A = coo_format(nrows, ncols, nnz, indx, jndx, rval)
   replaces user_module.f90:865-898 (coo_format) */
int idrs_coo_format_f64(int nrows, int ncols, int nnz, const int *indx, const int *jndx, const double *rval,
                        idrs_user_matrix_t *A);
/* coo_format(..., cval=), user_module.f90:865-898 */
int idrs_coo_format_c128(int nrows, int ncols, int nnz, const int *indx, const int *jndx, const double *cval,
                         idrs_user_matrix_t *A);
/* A = cds_format(nrows, ncols, half_band, rband) with rband(nrows,-hb:hb)
   replaces user_module.f90:932-958 (cds_format) */
int idrs_cds_format_f64(int nrows, int ncols, int half_band, const double *rband, idrs_user_matrix_t *A);
/* cds_format(..., cband=), user_module.f90:932-958 */
int idrs_cds_format_c128(int nrows, int ncols, int half_band, const double *cband, idrs_user_matrix_t *A);
/* Offset-list diagonal storage: values(nrows, ndiag) for strictly ascending
   offsets.  Extension of cds_format (user_module.f90:932-958) for wide
   sparse bands: the same per-row arithmetic over the stored offsets only. */
int idrs_dia_format_f64(int nrows, int ncols, int ndiag, const int *offsets, const double *vals,
                        idrs_user_matrix_t *A);
/* idrs_dia_format_f64 with complex values */
int idrs_dia_format_c128(int nrows, int ncols, int ndiag, const int *offsets, const double *vals,
                         idrs_user_matrix_t *A);
/* Matrix Market files (EXAMPLES/MATRIX_MARKET/mm_module.f90 over the NIST
   reader mmio.f): header of a file -- rep 1 coordinate / 2 array; field
   1 real / 2 complex / 3 integer / 4 pattern; symm 1 general / 2 symmetric
   / 3 skew-symmetric / 4 hermitian; entries: stored entries (coordinate)
   or rows*cols (array).  Replaces mminfo as mm_matrix calls it
   (mm_module.f90:85). */
int idrs_mm_info(const char *path, int *rep, int *field, int *symm, int *nrows, int *ncols, int64_t *entries);
/* A = mm_matrix(path, which_format) for a coordinate file,
   mm_module.f90:48-150 with mm_coo (:542-738) / mm_crs (:312-540): format
   3 'coo' or 2 'crs' (the same device CRS either way: the symmetric,
   skew-symmetric and hermitian halves are mirrored right after each
   off-diagonal entry, in file order); complex values pass through single
   precision as the reference's reader stores them (mm_module.f90:61). */
int idrs_mm_matrix(const char *path, int format, idrs_user_matrix_t *A);
/* the values of an array file (a right-hand-side block), column-major
   nrows x ncols into host vals ((re, im) pairs for complex files), the
   symmetric kinds expanded: mm_dense, mm_module.f90:152-310 */
int idrs_mm_dense(const char *path, double *vals);
/* Frees the device storage of a user matrix (the reference's allocatable
   components are released when the Fortran object goes out of scope). */
int idrs_user_matrix_destroy(idrs_user_matrix_t A);
/* Autotuner (no counterpart in the reference): times the CRS product with
   the lanes-per-row choices around the default and, when its padding is
   at most 1/4 of the entries, with a sliced-ELL copy of the matrix (one
   lane per row, *group = 1), reps products each, and keeps the fastest
   for all later products of K; *group receives it.  For one-thread-per-row
   products (DIA storage, the sliced copy) it then times 4096, 8192, 16384
   and 32768 row slabs (workgroups, at most one per 256 rows) and keeps the
   fastest.  A different choice
   changes the rounding order inside a row or of the epilogue's block sums,
   so the call is explicit; the sliced copy sums each row in CRS order like
   the reference's CRS product. */
int idrs_user_matrix_autotune(idrs_user_matrix_t K, int reps, int *group);
/* The product layout of K (no counterpart in the reference): lanes per row
   (1: sliced copy; DIA: 0) and row slabs of one-thread-per-row products. */
int idrs_user_matrix_layout(idrs_user_matrix_t K, int *group, int *row_slabs);
/* Sets the row slabs of K's one-thread-per-row products: 4096, 8192, 16384
   or 32768 (no counterpart in the reference; e.g. an earlier autotune's). */
int idrs_user_matrix_set_row_slabs(idrs_user_matrix_t K, int row_slabs);
/* Sets the lanes per row of K's CRS products (no counterpart in the
   reference): 1 (sliced copy, built here), 2, 4, 8, 16, 32 or 64 -- e.g. a
   choice an earlier autotune made.  Operators made on K pick the change up
   at their next solve or product call. */
int idrs_user_matrix_set_group(idrs_user_matrix_t K, int group);
/* nrows, ncols, stored entries, matrix_format (2 CRS, 3 COO, 4 CDS, 5 DIA) */
int idrs_user_matrix_info(idrs_user_matrix_t A, int *nrows, int *ncols, int64_t *nnz, int *format);

/* w = K*v: ruser_mv, user_module.f90:197-218 */
int idrs_user_mv_f64(idrs_user_matrix_t K, const double *v, double *w);
/* w = K*v: cuser_mv, user_module.f90:220-241 */
int idrs_user_mv_c128(idrs_user_matrix_t K, const double *v, double *w);
/* D = real_diagonal(K), user_module.f90:960-1047 (zeros replaced by 1) */
int idrs_real_diagonal(idrs_user_matrix_t K, double *D);
/* D = complex_diagonal(K), user_module.f90:1049-1161 */
int idrs_complex_diagonal(idrs_user_matrix_t K, double *D);

/* ---------------------------------------------------------------------
 * type(matrix): operator A = p(K D^-1) K D^-1 applied by operator(*)
 * ------------------------------------------------------------------- */

/* A%K = K (PP_IDRS/real_matrix.inc:9); K must outlive A */
int idrs_matrix_create(idrs_user_matrix_t K, idrs_matrix_t *A);
int idrs_matrix_destroy(idrs_matrix_t A);
/* A%P = REAL_PRECON(which_precon, K[, shift]): which_precon 0 (identity) or
   1 (diagonal scaling), user_module.f90:1327-1381; shift may be NULL. */
int idrs_matrix_set_precon_f64(idrs_matrix_t A, int which_precon, const double *shift);
/* A%P = COMPLEX_PRECON(which_precon, K[, shift]), user_module.f90:1465-1518;
   shift is one complex (re, im) or NULL. */
int idrs_matrix_set_precon_c128(idrs_matrix_t A, int which_precon, const double *shift);
/* Value entries one product of A reads (no reference counterpart: the
   reference always multiplies by every stored entry).  With diagonal
   scaling of a CDS matrix whose scaled main diagonal is exactly 1 that
   diagonal is not stored (*unit_diagonal = 1; NULL allowed) and the
   products are bit-identical to multiplying by the stored 1s. */
int idrs_matrix_stored_values(idrs_matrix_t A, int64_t *count, int *unit_diagonal);
/* A%rseed = seed (PP_IDRS/real_matrix.inc:41-45) */
int idrs_matrix_set_seed_f64(idrs_matrix_t A, double seed);
/* A%cseed = seed (complex_matrix.inc) */
int idrs_matrix_set_seed_c128(idrs_matrix_t A, const double *seed);
/* call real_preconditioner(A, preconditioner, degree, foci, center, H),
   PP_MATRIX/matrix_module.f90:551-635.  preconditioner: 0 none, 1 Neumann,
   2 Chebyshev, 3 Hessenberg.  foci (2 values), center (1 value) and
   H(ldh, ncol) may be NULL (absent). */
int idrs_real_preconditioner(idrs_matrix_t A, int preconditioner, int degree, const double *foci,
                             const double *center, const double *H, int ldh, int ncol);
/* call complex_preconditioner(...), matrix_module.f90:637-722; complex
   arguments interleaved. */
int idrs_complex_preconditioner(idrs_matrix_t A, int preconditioner, int degree, const double *foci,
                                const double *center, const double *H, int ldh, int ncol);
/* Sets A%degree and A%ralpha_p / A%calpha_p(0:degree) directly (values
   as (re, im) pairs); the polynomial of rpolpre, matrix_module.f90:266-304. */
int idrs_matrix_set_alpha(idrs_matrix_t A, int degree, const double *alpha);
/* alpha = chebyshev(degree, foci): rchebyshev / cchebyshev,
   matrix_module.f90:903-981; is_complex selects complex arithmetic, all
   values as (re, im) pairs (imaginary parts ignored for real). */
int idrs_chebyshev_coeffs(int degree, const double *foci, int is_complex, double *alpha);
/* alpha = neumann(degree, center): rneumann / cneumann, matrix_module.f90:851-901 */
int idrs_neumann_coeffs(int degree, const double *center, int is_complex, double *alpha);
/* alpha = compute_alpha(degree, H): rcompute_alpha / ccompute_alpha,
   matrix_module.f90:983-1057; H(ldh, ncol) as (re, im) pairs. */
int idrs_compute_alpha_coeffs(int degree, const double *H, int ldh, int ncol, int is_complex, double *alpha);
/* A%ralpha_p / A%calpha_p(0:degree) as (re, im) pairs; *degree out */
int idrs_matrix_get_alpha(idrs_matrix_t A, double *alpha, int maxlen, int *degree);
/* Generalized operators (PP_IDRS/real_matrix.inc, complex_matrix.inc set
   these components of type(matrix)).  Any of them sends A's products
   through the reference's general statements (matrix_module.f90:169-264):
   t = v/P, then K t with the seed terms K + seed C + seed^2 M (damping and
   mass), K + seed C + seed^2 I (damping), K - seed M (mass), or the block
   system [-C -K; I 0] - seed [M 0; 0 I] on vectors of 2 K%nrows.  Such
   operators run in double precision; M and C stay owned by the caller. */
/* A%M = M, A%mass_matrix = .true. (M NULL: none) */
int idrs_matrix_set_mass(idrs_matrix_t A, idrs_user_matrix_t M);
/* A%C = C, A%damping_matrix = .true. (C NULL: none) */
int idrs_matrix_set_damping(idrs_matrix_t A, idrs_user_matrix_t C);
/* A%block_system, A%multishift (the latter selects the deflation
   projector I - V V^H, matrix_module.f90:1376-1383) */
int idrs_matrix_set_flags(idrs_matrix_t A, int block_system, int multishift);
/* rows of A's vectors: 2 K%nrows for a block system, else K%nrows */
int idrs_matrix_rows(idrs_matrix_t A, int *rows);
/* A%P = REAL_PRECON / COMPLEX_PRECON(1, PK, PM, PC, shift)
   (user_module.f90:1327-1381, 1465-1530): v/A divides by
   diag(PK) [+ shift diag(PC) + shift^2 diag(PM) | + shift diag(PC) + shift^2
   | - shift diag(PM) | - shift]; PM, PC, shift may be NULL; shift as one
   value or an (re, im) pair when is_complex.  E.g. a lumped mass matrix:
   PK = M. */
int idrs_matrix_set_precon_diag(idrs_matrix_t A, idrs_user_matrix_t PK, idrs_user_matrix_t PM, idrs_user_matrix_t PC,
                                const double *shift, int is_complex);
/* A%P = REAL_PRECON / COMPLEX_PRECON(which, PK, PM, PC, shift),
   user_module.f90:1383-1459 (real), 1531-1602 (complex): which 2 dense LU,
   3 band LU (half bandwidth from get_band, :1833-1872) of
   PK [+ shift PC + shift^2 PM | + shift PC + shift^2 I | - shift PM | - shift I]
   (PK alone without a shift), factored on the device without pivoting
   (:1604-1701, :2068-2179); v/A then solves with the factors (:1703-1831,
   :2181-2336).  shift: NULL, one value, or an (re, im) pair when is_complex;
   complex factors when is_complex or any matrix is complex.  A zero pivot
   returns an error ("Pivot equal to zero in LU decomposition").  Dense LU
   takes CRS/COO matrices. */
int idrs_matrix_set_precon_lu(idrs_matrix_t A, int which, idrs_user_matrix_t PK, idrs_user_matrix_t PM,
                              idrs_user_matrix_t PC, const double *shift, int is_complex);
/* A's preconditioner = B's LU factors, shared (no second factorization) */
int idrs_matrix_share_lu(idrs_matrix_t A, idrs_matrix_t B);
/* A%shift_invert and A%rshift / A%cshift (matrix_module.f90:306-549,
   724-806): the block system's preconditioner inverts [-C-sM -K; I -sI]
   through A%P; ms scale-back multiplies by (seed - shift)/(sigma - shift);
   the multishift deflation step uses (sigma - seed)/(sigma - shift) */
int idrs_matrix_set_shift_invert(idrs_matrix_t A, int shift_invert, const double *shift, int is_complex);
/* make_deflation(A, V), matrix_module.f90:1284-1352: V(rows, nd) host,
   nd <= 16; AV = real_mv(A, V) and V^H AV are formed on the device */
int idrs_make_deflation_f64(idrs_matrix_t A, int nd, const double *V);
int idrs_make_deflation_c128(idrs_matrix_t A, int nd, const double *V);
/* w = Pmul(A, v), matrix_module.f90:1369-1431 (the real version subtracts
   matmul(A%rV, t), the complex one matmul(A%cAV, t), as written) */
int idrs_pmul_f64(idrs_matrix_t A, const double *v, double *w);
int idrs_pmul_c128(idrs_matrix_t A, const double *v, double *w);
/* w = real_mv(A, v) / complex_mv(A, v) alone (matrix_module.f90:169-264) */
int idrs_system_mv_f64(idrs_matrix_t A, const double *v, double *w);
int idrs_system_mv_c128(idrs_matrix_t A, const double *v, double *w);
/* x = add_deflation(A, b, y), matrix_module.f90:1433-1506 */
int idrs_add_deflation_f64(idrs_matrix_t A, const double *b, const double *y, double *x);
int idrs_add_deflation_c128(idrs_matrix_t A, const double *b, const double *y, double *x);
/* x = scaleback(A, b, y) with its deflation step, matrix_module.f90:808-849 */
int idrs_scaleback_b_f64(idrs_matrix_t A, const double *b, const double *y, double *x);
int idrs_scaleback_b_c128(idrs_matrix_t A, const double *b, const double *y, double *x);
/* X = add_deflation(A, sigma, b, Y), rms_add_deflation /
   cms_add_deflation (matrix_module.f90:1508-1623): per shift k,
   V Hs^-1 V^H b + (y_k - V Hs^-1 V^H (real_mv(A, y_k) - (sigma_k - seed) y_k)),
   Hs = V^H AV - (sigma_k - seed) I; Y, X (rows x n_sigma) column-major host
   arrays, sigma as for idrs_ms_scaleback_b_*; Y copied without deflation */
int idrs_ms_add_deflation_f64(idrs_matrix_t A, int n_sigma, const double *sigma, const double *b, const double *Y,
                              double *X);
int idrs_ms_add_deflation_c128(idrs_matrix_t A, int n_sigma, const double *sigma, const double *b, const double *Y,
                               double *X);
/* X = scaleback(A, sigma, b, Y), matrix_module.f90:724-806 without
   shift-and-invert: per shift k, p_k(A) y_k (alpha_s as for
   idrs_ms_scaleback_*), ms_add_deflation for sigma_k (:1508-1623), x/A;
   sigma(n_sigma) real, or (re, im) pairs for c128; host arrays */
int idrs_ms_scaleback_b_f64(idrs_matrix_t A, int n_sigma, const double *sigma, int degree, const double *alpha_s,
                            const double *b, const double *Y, double *X);
int idrs_ms_scaleback_b_c128(idrs_matrix_t A, int n_sigma, const double *sigma, int degree, const double *alpha_s,
                             const double *b, const double *Y, double *X);
/* w = A*v: rmatvec, matrix_module.f90:152-167 */
int idrs_matvec_f64(idrs_matrix_t A, const double *v, double *w);
/* w = A*v: cmatvec, matrix_module.f90:135-150 */
int idrs_matvec_c128(idrs_matrix_t A, const double *v, double *w);
/* W(:,j) = A*V(:,j), j = 1..ncols (V, W: n x ncols, column-major): rmatvec /
   cmatvec (matrix_module.f90:135-167) over a block of vectors, each column
   bit-identical to idrs_matvec_* on it.  No single counterpart in the
   reference, whose multishift solvers apply A once per shift
   (msidrs.inc:312-320): here up to 8 columns share each product's read of
   the matrix.  A must be square. */
int idrs_matvec_block_f64(idrs_matrix_t A, int ncols, const double *V, double *W);
int idrs_matvec_block_c128(idrs_matrix_t A, int ncols, const double *V, double *W);
/* w = v/A: rprecon, matrix_module.f90:306-352 */
int idrs_precon_f64(idrs_matrix_t A, const double *v, double *w);
/* w = v/A: cprecon, matrix_module.f90:354-400 */
int idrs_precon_c128(idrs_matrix_t A, const double *v, double *w);
/* x = scaleback(A, b, y): rscaleback, matrix_module.f90:831-849 */
int idrs_scaleback_f64(idrs_matrix_t A, const double *y, double *x);
/* x = scaleback(A, b, y): cscaleback, matrix_module.f90:808-829 */
int idrs_scaleback_c128(idrs_matrix_t A, const double *y, double *x);
/* X(:,k) = ms_scaleback(A, sigma, b, Y)(:,k) for the shifted polynomials:
   X(:,k) = p_k(K') Y(:,k) / D with alpha_s(0:degree, k) as (re, im) pairs,
   column-major (degree+1) x n_sigma (MS_PARAMETERS output);
   ms_rscaleback / ms_cscaleback, matrix_module.f90:724-806 (no deflation,
   no shift-and-invert, as elsewhere in this library).  Y and X are device
   pointers if device_pointers != 0; alpha_s is always a host array. */
int idrs_ms_scaleback_f64(idrs_matrix_t A, int n_sigma, int degree, const double *alpha_s, const double *Y,
                          double *X, int device_pointers);
int idrs_ms_scaleback_c128(idrs_matrix_t A, int n_sigma, int degree, const double *alpha_s, const double *Y,
                           double *X, int device_pointers);

/* ---------------------------------------------------------------------
 * solvers
 * ------------------------------------------------------------------- */

/* Optional arguments of IDRS (MODULES/IDRS/idrs_module.f90:180-251).  has_*
   flags play the role of Fortran's present(). */
typedef struct {
    int32_t has_M1;             /* present(M1)                            */
    int32_t has_tolerance;
    int32_t has_maximum_iterations;
    int32_t has_variant;
    idrs_matrix_t M1;           /* v/M1 divides by M1's diagonal if set   */
    double tolerance;
    int32_t maximum_iterations;
    int32_t variant;            /* 1 IDR(s), 2 BiCGSTAB, 3/4 one-image minsync */
    const void *x0;             /* x0(n)                                  */
    const void *U0;             /* U0(n,s)                                */
    const void *omega;          /* omega(n_omega), host                   */
    int32_t n_omega;
    int32_t resvec_len;
    double *resvec;             /* resvec(maxit+1), host, output          */
    const void *PT_raw;         /* raw shadow space PT(s,n) before Gram-Schmidt;
                                   NULL: the library's uniform stream      */
    uint64_t p_seed;            /* seed of that stream (0 -> 1, the reference's this_image()) */
    int32_t device_pointers;    /* b, x, x0, U0, PT_raw are device pointers */
    int32_t time_spmv;          /* time each SpMV launch with HIP events  */
    int32_t graph_mode;         /* 0: replay each IDR(s) cycle from a HIP graph
                                   when n <= 2^21 (launch-bound sizes; the
                                   batched solve always); 1: always;
                                   2: never.  No reference counterpart; the
                                   results are identical either way        */
    int32_t H_nritz;            /* H(H_nritz+1, H_nritz): size(H,1)-1      */
    void *H;                    /* IDRS only: Hessenberg matrix, host,
                                   column-major, output (idrs.inc:81-90,
                                   179-181, 211-213, 231-237); NULL: absent */
} idrs_opts_t;

typedef struct {
    int32_t flag;               /* 0 converged, 1 maxit, 2 true residual above tol, 3 breakdown */
    int32_t iterations;
    double relres;              /* ||b - A x|| / ||b|| (true residual)   */
    double solve_ms;            /* device time of the solve (HIP events)  */
    double spmv_ms;             /* summed SpMV launch time (time_spmv)    */
    int32_t spmv_launches;
    int32_t reserved;
} idrs_stats_t;

/* x = IDRS(A, b, s, [M1, tolerance, maximum_iterations, variant, flag,
   relres, iterations, x0, U0, omega, resvec, H]): RIDRS,
   MODULES/IDRS/idrs_module.f90:180-251 with MODULES/IDRS/idrs.inc:1-311. */
int idrs_idrs_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                  idrs_stats_t *stats);
/* CIDRS, idrs_module.f90:107-178 */
int idrs_idrs_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                   idrs_stats_t *stats);
/* x = IDRS_RITZ(A, b, s, [M, tolerance, maximum_iterations, variant,
   flag, relres, iterations, x0, U0, omega, resvec, H]): RIDRS_RITZ,
   MODULES/IDRS/ritz_module.f90:427-556.  opts as for idrs_idrs_f64, with
   has_M1/M1 the argument M (IDRS's preconditioner and v/M in compute_U0);
   opts' U0, omega and H must be absent (they are outputs here).  Outputs
   (NULL / 0: absent): omega(n_omega), host; U0(n, dim_u0) -- a device
   pointer when opts->device_pointers, else host; H(h_rows, h_cols) =
   Hf(1:h_rows, 1:h_cols) of the solve's Hf(maxit+1, maxit), host;
   *nritz = (iter/(s+1)) s + mod(iter, s+1); *converged = Ritz values the
   QR iteration converged for the omegas.  The reference's stop statements
   ("Too many parameters omega requested, reduce omega.", "Recycle space
   is too large, reduce dimension of U0.", "Dimensions of Hessenberg
   matrix too large") return IDRS_ERR_ARG with that message.  The
   shift-and-invert power method starts from the library's uniform stream
   (seed p_seed) after the shadow space's values; the reference continues
   its compiler's random_number stream instead. */
int idrs_idrs_ritz_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                       idrs_stats_t *stats, double *omega, int n_omega, double *U0, int dim_u0, double *H, int h_rows,
                       int h_cols, int *nritz, int *converged);
/* CIDRS_RITZ, ritz_module.f90:225-425 (complex arrays as (re, im) pairs) */
int idrs_idrs_ritz_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                        idrs_stats_t *stats, double *omega, int n_omega, double *U0, int dim_u0, double *H,
                        int h_rows, int h_cols, int *nritz, int *converged);
/* Spectral routines on small dense matrices (host, no GPU).  Matrices
   column-major with leading dimension ldh >= n; complex values as (re, im)
   pairs, H too (imaginary parts ignored unless is_complex).
   eig(n) = QR_eig(H(1:n,1:n)) (dense_la_module.f90:192-403: CQR_eig for
   is_complex, else RQR_eig), U(n,n) its Schur vectors when non-NULL,
   *converged when non-NULL. */
int idrs_qr_eig(int n, const double *H, int ldh, int is_complex, double *eig, double *U, int *converged);
/* outlier(n) = filter_ritzval(ritzval(n)) (matrix_module.f90:1059-1077), 0/1 */
int idrs_filter_ritzval(int n, const double *ritzval, int *outlier);
/* analyse_spectrum(H(1:n,1:n), foci, center, shift)
   (matrix_module.f90:1079-1282); each output (2, 1, 1 values) may be NULL */
int idrs_analyse_spectrum(int n, const double *H, int ldh, int is_complex, double *foci, double *center,
                          double *shift);
/* schurvec(n, s) = SI_power(H(1:n,1:n), s, shift) (dense_la_module.f90:
   455-603); shift NULL: 0; start block from the uniform stream (seed) at
   counter offset */
int idrs_si_power(int n, const double *H, int ldh, int is_complex, int s, const double *shift, uint64_t seed,
                  uint64_t offset, double *schurvec);
/* idrs_idrs_f64 that allocates nothing.  No counterpart in the reference:
   for callers that cannot allocate device or pinned memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   b, x and opts' x0, U0, PT_raw are device pointers (device_pointers is
   implied); every workspace buffer comes from a fixed 64 MiB region in the
   library's device data (one per GPU, held until the process ends), so a
   system whose workspace is larger -- about (3s + 7) n doubles plus the
   reduction partials, e.g. n up to ~450 000 rows at s = 4 -- is refused
   with IDRS_ERR_ARG before anything runs.  The solve runs eagerly (no
   graph), reads the convergence flag after every cycle with a blocking
   copy (no events, no pinned memory) and reports host-clock solve_ms; SpMV
   timing (time_spmv) is rejected.  An operator whose K uses the sliced copy
   must have had one ordinary product or solve first (that builds its
   sliced values).  Results are bit-identical to idrs_idrs_f64's.  All
   calls on one GPU share the region: calls that overlap on one GPU (other
   host threads, other streams) race. */
int idrs_idrs_f64_noalloc(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                          idrs_stats_t *stats);

/* Optional arguments of QMRIDR, MSIDRS and MSQMRIDR
   (MODULES/IDRS/idrs_module.f90:253-713).  has_* flags are present(). */
typedef struct {
    int32_t has_M1;
    int32_t has_tolerance;
    int32_t has_maximum_iterations;
    int32_t has_variant;        /* MSIDRS: 1 msidrs, 2 msbicgstab         */
    idrs_matrix_t M1;           /* QMRIDR: preconditioner; MS*: mass matrix */
    double tolerance;
    int32_t maximum_iterations;
    int32_t variant;
    int32_t has_inner_s;        /* QMRIDR, MSQMRIDR                       */
    int32_t inner_s;
    int32_t has_inner_tolerance;
    int32_t has_inner_iterations;
    double inner_tolerance;
    int32_t inner_iterations;
    int32_t n_omega;
    const void *x0;             /* QMRIDR: x0(n)                          */
    const void *omega;          /* omega(n_omega), host                   */
    double *resvec;             /* resvec(resvec_len, n_sigma), host, output */
    int32_t resvec_len;         /* leading dimension of resvec, >= maxit+1 */
    int32_t device_pointers;    /* b, x, x0, PT_raw are device pointers   */
    const void *PT_raw;         /* raw shadow space PT(s,n); NULL: uniform stream */
    uint64_t p_seed;
    void *colfac;               /* MSIDRS: colfac(n_sigma), host, output  */
} idrs_krylov_opts_t;

/* x = QMRIDR(A, b, s, [M1, tolerance, maximum_iterations, flag, relres,
   iterations, inner_s, inner_tolerance, inner_iterations, x0, omega,
   resvec]): RQMRIDR, idrs_module.f90:331-408 with qmridr.inc:1-263. */
int idrs_qmridr_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_krylov_opts_t *opts, double *x,
                    idrs_stats_t *stats);
/* CQMRIDR, idrs_module.f90:253-329 */
int idrs_qmridr_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_krylov_opts_t *opts, double *x,
                     idrs_stats_t *stats);
/* X(n,n_sigma) = MSIDRS(A, b, sigma, s, [M1, tolerance,
   maximum_iterations, variant, flag, relres(n_sigma), iterations, omega,
   resvec, colfac]): RMSIDRS, idrs_module.f90:572-641 with msidrs.inc:1-334.
   sigma is a host array (complex: interleaved pairs); relres may be NULL;
   stats->relres is the largest of relres. */
int idrs_msidrs_f64(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                    const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats);
/* CMSIDRS, idrs_module.f90:643-713 */
int idrs_msidrs_c128(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                     const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats);
/* X(n,n_sigma) = MSQMRIDR(A, b, sigma, s, [M1, tolerance,
   maximum_iterations, flag, relres(n_sigma), iterations, inner_s,
   inner_tolerance, inner_iterations, omega, resvec]): RMSQMRIDR,
   idrs_module.f90:491-570 with msqmridr.inc:1-293. */
int idrs_msqmridr_f64(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                      const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats);
/* CMSQMRIDR, idrs_module.f90:410-489 */
int idrs_msqmridr_c128(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                       const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats);

/* ---------------------------------------------------------------------
 * mixed precision (no counterpart in the reference, which is FP64 only)
 * ------------------------------------------------------------------- */

/* x = IDRS(A, b, s, ...) to the FP64 tolerance opts->tolerance with
   single-precision inner solves: r = b - A x in FP64, then IDR(s) on
   A d = r/||r|| to inner_tolerance whose SpMVs read FP32 copies of the
   (scaled) matrix values -- with FP32 vectors too if fp32_vectors != 0,
   FP64 vectors otherwise -- then x += ||r|| d, repeated (at most max_outer
   times) until the true FP64 residual meets the tolerance.
   opts->maximum_iterations bounds the inner iterations in total; x0,
   PT_raw, variant and device_pointers as for idrs_idrs_f64 (M1, U0, omega,
   resvec rejected).  Matrix values must be real and the operator without a
   seed shift.  stats->iterations: inner iterations; *outer_steps: FP64
   residual replacements. */
int idrs_idrs_mixed_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts,
                        double inner_tolerance, int max_outer, int fp32_vectors, double *x, idrs_stats_t *stats,
                        int *outer_steps);
/* complex128 vectors on a real-valued operator; fp32_vectors must be set
   (complex64 inner solves) */
int idrs_idrs_mixed_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts,
                         double inner_tolerance, int max_outer, int fp32_vectors, double *x, idrs_stats_t *stats,
                         int *outer_steps);

/* X(:,k) = (A - sigma_k I)^-1 b for every shift, the multishift solvers in
   complex64 with FP64 refinement: MSQMRIDR (method 0) or MSIDRS (method 1)
   with complex64 vectors and FP32 matrix values, every shift at once, to
   inner_tolerance; then each shift's FP64 residual r_k = b - (A - sigma_k I)
   x_k, and for a shift above opts->tolerance complex64 corrections
   (A - sigma_k I) d = r_k/||r_k|| by the same solver over that shift,
   x_k += ||r_k|| d (at most max_outer rounds).  Host scalars (the small
   recurrences, omegas, norms) stay FP64.  sigma: (re, im) pairs; X (n,
   n_sigma) column-major; relres (host) each shift's FP64 true residual;
   stats->iterations: all complex64 iterations; *outer_steps: FP64 residual
   rounds; *corrections: single-shift correction solves.  s, variant, inner
   solver and device_pointers as for idrs_msqmridr_c128 / idrs_msidrs_c128
   (M1, x0, omega, resvec, colfac and PT_raw rejected); real matrix values,
   no seed shift. */
int idrs_ms_mixed_c128(idrs_matrix_t A, int method, int n, const double *b, int n_sigma, const double *sigma, int s,
                       const idrs_krylov_opts_t *opts, double inner_tolerance, int max_outer, double *X,
                       double *relres, idrs_stats_t *stats, int *outer_steps, int *corrections);

/* ---------------------------------------------------------------------
 * differentiable solves (no counterpart in the reference): the adjoint
 * system A^T lambda = dL/dx is solved by IDRS on the transposed matrix, and
 * dL/dA on A's sparsity pattern is -lambda_i x_j.
 * ------------------------------------------------------------------- */

/* KT = transpose of a CRS/COO user matrix (CRS; keeps the value map) */
int idrs_crs_transpose(idrs_user_matrix_t K, idrs_user_matrix_t *KT);
/* KT's values = K's values after K's values changed */
int idrs_crs_transpose_refresh(idrs_user_matrix_t K, idrs_user_matrix_t KT);
/* replaces the nnz stored values of a CRS matrix (same pattern, CRS entry
   order); vals is a device pointer if device_pointer != 0.  The diagonal
   is recomputed; operators built with diagonal scaling or low-precision
   copies on K must be rebuilt. */
int idrs_user_matrix_set_values_f64(idrs_user_matrix_t K, const double *vals, int device_pointer);
/* grad[k] = -lambda[i] * x[j] for every stored entry k = (i, j) of the CRS
   matrix K, in CRS entry order (device pointers if device_pointers != 0) */
int idrs_crs_pattern_grad_f64(idrs_user_matrix_t K, const double *lambda, const double *x, double *grad,
                              int device_pointers);

/* ---------------------------------------------------------------------
 * batches of independent systems (no counterpart in the reference, which
 * solves one system per call; these run IDRS, idrs.inc:1-311, on every
 * system of a batch at once)
 * ------------------------------------------------------------------- */

/* A = block-diagonal matrix of nsys square CRS/COO user matrices (copies);
   vectors of a batch are the systems' vectors concatenated in order. */
int idrs_batch_format(int nsys, const idrs_user_matrix_t *K, idrs_user_matrix_t *A);
/* nsys of a batch matrix and, if row0 != NULL, the first row of every
   system (row0[nsys] = total rows); nsys = 0 for an ordinary matrix. */
int idrs_batch_info(idrs_user_matrix_t A, int *nsys, int64_t *row0);
/* x = IDRS(A_k, b_k, s, ...) for every system k of the batch matrix behind
   A (created from a batch user matrix; diagonal scaling and polynomial
   preconditioners apply to every system).  Options as idrs_idrs_f64 except
   M1, U0 and resvec (rejected); without maximum_iterations each system
   uses min(2 n_k, 1000).  stats: one per system (solve_ms: whole batch). */
int idrs_idrs_batched_f64(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double *x,
                          idrs_stats_t *stats);
/* x_k = IDRS(A_k, b_k, s, ...) for every system k of the batch matrix
   behind A to the FP64 tolerance opts->tolerance on each system's true
   residual, in mixed precision: r_k = b_k - A_k x_k in FP64, then batched
   IDR(s) on A d_k = r_k/|r_k| whose SpMVs read FP32 copies of the matrix
   values (vectors FP64) to inner_tolerance (no tighter than the tolerance
   needs), x_k += |r_k| d_k, at most max_outer times per system.
   opts->maximum_iterations bounds each inner solve (default min(2 n_max,
   1000)).  Real matrix values, no seed shift; M1, U0, omega, resvec, x0 and
   PT_raw are rejected.  stats and outer_steps: one per system
   (stats->iterations: inner iterations; solve_ms: whole batch). */
int idrs_idrs_batched_mixed_f64(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts,
                                double inner_tolerance, int max_outer, double *x, idrs_stats_t *stats,
                                int *outer_steps);
/* complex128 form of idrs_idrs_batched_f64 */
int idrs_idrs_batched_c128(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double *x,
                           idrs_stats_t *stats);

/* PT = MAKE_P(...) for inspection: the orthonormalised shadow space PT(s,n)
   (RMAKE_P / CMAKE_P, idrs_module.f90:839-937).  method 1 Gram-Schmidt of
   PT_raw (or the library stream when NULL); method 2 PT = conj(r). Host arrays. */
int idrs_make_p_f64(int n, int s, int method, const double *PT_raw, const double *r, uint64_t seed, double *PT);
int idrs_make_p_c128(int n, int s, int method, const double *PT_raw, const double *r, uint64_t seed, double *PT);

/* Number of HIP devices visible (0 without a GPU); never fails. */
int idrs_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* IDRS_CAPI_H */
