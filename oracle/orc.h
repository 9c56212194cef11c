/*
 * Oracle: a plain-C restatement of the reference IDRS
 * package's hot path (IDR(s) solvers, operator, polynomial preconditioner,
 * sparse formats).  TEST INFRASTRUCTURE ONLY: it is the checker that the
 * HIP library is compared against (liborc.so, single-threaded) and, built
 * with OpenMP row and vector loops (liborc_omp.so), the CPU baseline timed
 * by bench.py.
 * Nothing in the product links, loads or calls it.
 *
 * Every routine names the reference file:line it restates (paths relative to
 * the reference's IDRS/ directory).  Arithmetic follows the reference's
 * statement order: one vector statement at a time, sequential sums, the same
 * single-precision literals (kappa = 0.7 -> 0.699999988079071, sqrt(2.) in
 * single precision).  The build uses -ffp-contract=off so every multiply and
 * add rounds separately, as the reference's array statements do.
 *
 * Indexing: all index arrays are 1-based (Fortran convention), vectors are
 * 0-based C arrays of length n.  Column-major (Fortran) layout for 2-D arrays:
 * PT(s,n) -> PT[i + s*j]; G(n,s) -> G[j + n*i]; band(n,-hb:hb) ->
 * band[i + n*(d+hb)].
 */
#ifndef ORC_H
#define ORC_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex orc_cplx;

/* threads of the parallel loops (OpenMP build); returns the count in use */
int orc_threads(int n);

/* user_matrix formats, PP_MATRIX/user_module.f90:134-167 */
enum { ORC_CRS = 2, ORC_COO = 3, ORC_CDS = 4, ORC_DIA = 5 };

typedef struct {
    int format;
    int nrows, ncols;
    /* values: exactly one of rval / cval set (or neither: pattern matrix) */
    const double *rval;
    const orc_cplx *cval;
    /* CRS (user_module.f90:829-863) */
    const int *row_ptr, *col_ind;
    /* COO (user_module.f90:865-898) */
    int nnz;
    const int *indx, *jndx;
    /* CDS dense band (user_module.f90:932-958): values in rval/cval,
       laid out (nrows, -half_band:half_band) column-major */
    int half_band;
    /* DIA offset list (extension; same arithmetic as CDS restricted to the
       stored offsets, in ascending offset order): values (nrows, ndiag) */
    int ndiag;
    const int *offsets;
} orc_user_matrix;

/* type(matrix), PP_MATRIX/matrix_module.f90:38-81: K, the diagonal
   preconditioner P (which_precon 0/1), seed shift, polynomial coefficients,
   mass and damping matrices, block systems and deflation. */
typedef struct {
    orc_user_matrix K;
    int which_precon;          /* 0 identity, 1 diagonal scaling            */
    const double *rdiag;       /* P%rband(:,0) when which_precon == 1 (real) */
    const orc_cplx *cdiag;     /* P%cband(:,0) (complex)                     */
    double rseed;              /* A%rseed                                   */
    orc_cplx cseed;            /* A%cseed                                   */
    int degree;                /* A%degree                                  */
    const double *ralpha_p;    /* A%ralpha_p(0:degree)                      */
    const orc_cplx *calpha_p;  /* A%calpha_p(0:degree)                      */
    const orc_user_matrix *M;  /* A%M when A%mass_matrix, else NULL         */
    const orc_user_matrix *C;  /* A%C when A%damping_matrix, else NULL      */
    int block_system;          /* A%block_system: vectors of 2 K%nrows      */
    int multishift;            /* A%multishift (deflation projector I - VV^H) */
    int n_deflation;           /* A%n_deflation                             */
    void *V, *AV, *VtAV;       /* A%rV/cV (n,nd), A%rAV/cAV (n,nd), A%rVtAV/cVtAV
                                  (nd,nd), T column-major; AV and VtAV are
                                  filled by make_deflation                  */
    /* which_precon 2 / 3: P%rmat/cmat (n,n) or P%rband/cband (n,-hb:hb),
       column-major T, holding the LU factors made by orc_*_precon_lu */
    const void *lu;
    int lu_hb;
    int shift_invert;          /* A%shift_invert                            */
    double rshift;             /* A%rshift                                  */
    orc_cplx cshift;           /* A%cshift                                  */
} orc_matrix;

/* optional arguments of IDRS (idrs_module.f90:180-251).  has_* flags mirror
   Fortran's present(). */
typedef struct {
    int has_M1;  const orc_matrix *M1;
    int has_tol; double tol;
    int has_maxit; int maxit;
    int has_variant; int variant;
    const void *x0;             /* T[n] or NULL                     */
    const void *U0;             /* T[n*s] column-major or NULL      */
    const void *omega; int n_omega;  /* T[n_omega] or NULL          */
    double *resvec; int resvec_len;  /* out, or NULL                */
    /* Shadow space: raw random PT(s,n) before orthonormalisation.  NULL ->
       the oracle's default uniform [0,1) stream (orc_uniform01). */
    const void *PT_raw;
    /* QMRIDR / MSQMRIDR inner solver parameters (present() flags) */
    int has_inner_s; int inner_s;
    int has_inner_tol; double inner_tol;
    int has_inner_it; int inner_it;
    /* multishift outputs: relres(n_sigma), colfac(n_sigma) (T), or NULL;
       resvec is (resvec_len, n_sigma) column-major for the multishift solvers */
    double *relres_vec;
    void *colfac;
    /* IDRS: Hessenberg matrix H(H_nritz+1, H_nritz), T column-major, output
       (idrs.inc:81-90, 179-181, 211-213, 231-237), or NULL */
    void *H; int H_nritz;
} orc_idrs_opts;

typedef struct {
    int flag;
    double relres;
    int iterations;
} orc_stats;

/* ---- default shadow-space stream (uniform [0,1), counter based) ---- */
double orc_uniform01(uint64_t seed, uint64_t index);

/* ---- user_module restatement ---- */
void orc_r_user_mv(const orc_user_matrix *A, const double *v, double *w);
void orc_c_user_mv(const orc_user_matrix *A, const orc_cplx *v, orc_cplx *w);
void orc_r_diagonal(const orc_user_matrix *A, double *D);
void orc_c_diagonal(const orc_user_matrix *A, orc_cplx *D);

/* ---- matrix_module restatement ---- */
void orc_r_matvec(const orc_matrix *A, const double *v, double *w);
void orc_c_matvec(const orc_matrix *A, const orc_cplx *v, orc_cplx *w);
void orc_r_precon(const orc_matrix *A, const double *v, double *w);
void orc_c_precon(const orc_matrix *A, const orc_cplx *v, orc_cplx *w);
void orc_r_scaleback(const orc_matrix *A, const double *y, double *x);
void orc_c_scaleback(const orc_matrix *A, const orc_cplx *y, orc_cplx *x);
/* rows of the operator's vectors (2 K%nrows for a block system) */
int  orc_r_rows(const orc_matrix *A);
int  orc_c_rows(const orc_matrix *A);
/* real_mv / complex_mv alone (matrix_module.f90:169-264) */
void orc_r_system_mv(const orc_matrix *A, const double *v, double *w);
void orc_c_system_mv(const orc_matrix *A, const orc_cplx *v, orc_cplx *w);
/* Pmul in place (matrix_module.f90:1369-1431) */
void orc_r_pmul(const orc_matrix *A, double *w);
void orc_c_pmul(const orc_matrix *A, orc_cplx *w);
/* make_deflation (matrix_module.f90:1284-1352): fills A->AV, A->VtAV from A->V */
void orc_r_make_deflation(orc_matrix *A);
void orc_c_make_deflation(orc_matrix *A);
/* add_deflation (matrix_module.f90:1433-1506) */
void orc_r_add_deflation(const orc_matrix *A, const double *b, const double *y, double *x);
void orc_c_add_deflation(const orc_matrix *A, const orc_cplx *b, const orc_cplx *y, orc_cplx *x);
/* rms_add_deflation / cms_add_deflation (matrix_module.f90:1508-1623) for
   shifts sigma (n_sigma), Y and X (rows, n_sigma) column-major */
void orc_r_ms_add_deflation(const orc_matrix *A, int n_sigma, const double *sigma, const double *b,
                            const double *Y, double *X);
void orc_c_ms_add_deflation(const orc_matrix *A, int n_sigma, const orc_cplx *sigma, const orc_cplx *b,
                            const orc_cplx *Y, orc_cplx *X);
/* rscaleback / cscaleback with the deflation step (matrix_module.f90:808-849) */
void orc_r_scaleback_b(const orc_matrix *A, const double *b, const double *y, double *x);
void orc_c_scaleback_b(const orc_matrix *A, const orc_cplx *b, const orc_cplx *y, orc_cplx *x);
int  orc_r_compute_alpha(int degree, const double *H, int ldh, int ncol, double *alpha);
int  orc_c_compute_alpha(int degree, const orc_cplx *H, int ldh, int ncol, orc_cplx *alpha);
void orc_r_chebyshev(int degree, const double foci[2], double *alpha);
void orc_c_chebyshev(int degree, const orc_cplx foci[2], orc_cplx *alpha);
void orc_r_neumann(int degree, double center, double *alpha);
void orc_c_neumann(int degree, orc_cplx center, orc_cplx *alpha);
/* REAL_PRECON / COMPLEX_PRECON(2|3, K, M, C, shift) (user_module.f90:
   1383-1459, 1531-1602): the dense (which 2) or band (which 3) matrix
   K [+ shift C + shift^2 M | + shift C + shift^2 I | - shift M | - shift I]
   (shift NULL: K alone), then RLU_decomp / band_LU_decomp
   (:1604-1701, :2068-2179), in place.  lu: (n,n) or (n,-hb:hb) column-major,
   hb from get_band (:1833-1872) -- orc_*_get_band gives it for sizing.
   Returns 0, or -1 on a zero pivot (the reference's error stop). */
int  orc_get_band(const orc_user_matrix *K, const orc_user_matrix *M, const orc_user_matrix *C);
int  orc_r_precon_lu(int which, const orc_user_matrix *K, const orc_user_matrix *M, const orc_user_matrix *C,
                     const double *shift, int hb, double *lu);
int  orc_c_precon_lu(int which, const orc_user_matrix *K, const orc_user_matrix *M, const orc_user_matrix *C,
                     const orc_cplx *shift, int hb, orc_cplx *lu);
/* RLU_solve / band_LU_solve (user_module.f90:1703-1831, 2181-2336) */
void orc_r_lu_solve(int which, int n, int hb, const double *lu, const double *b, double *x);
void orc_c_lu_solve(int which, int n, int hb, const orc_cplx *lu, const orc_cplx *b, orc_cplx *x);
/* ms_rscaleback / ms_cscaleback (matrix_module.f90:724-806): per shift the
   polynomial alpha_s(:,k) (degree A->degree, or y itself), the deflation
   step, the shift-and-invert factor (seed-shift)/(sigma-shift), then /A;
   Y, X (rows, n_sigma) column-major */
void orc_r_ms_scaleback_b(const orc_matrix *A, int n_sigma, const double *sigma, const double *alpha_s,
                          const double *b, const double *Y, double *X);
void orc_c_ms_scaleback_b(const orc_matrix *A, int n_sigma, const orc_cplx *sigma, const orc_cplx *alpha_s,
                          const orc_cplx *b, const orc_cplx *Y, orc_cplx *X);
int  orc_r_ms_parameters(const orc_matrix *A, int n_sigma, const double *sigma,
                         double *sigma_p, double *alpha_s);
int  orc_c_ms_parameters(const orc_matrix *A, int n_sigma, const orc_cplx *sigma,
                         orc_cplx *sigma_p, orc_cplx *alpha_s);

/* ---- idrs_module restatement ---- */
double orc_r_norm(int n, const double *w);
double orc_c_norm(int n, const orc_cplx *w);
void orc_r_make_p(int n, int s, int method, const double *PT_raw, const double *r, double *PT);
void orc_c_make_p(int n, int s, int method, const orc_cplx *PT_raw, const orc_cplx *r, orc_cplx *PT);
int orc_r_idrs(const orc_matrix *A, int n, const double *b, int s,
               const orc_idrs_opts *o, double *x, orc_stats *st);
int orc_c_idrs(const orc_matrix *A, int n, const orc_cplx *b, int s,
               const orc_idrs_opts *o, orc_cplx *x, orc_stats *st);

/* ---- dense_la_module restatement ---- */
void orc_r_solve(int n, const double *A, const double *b, double *x);
void orc_c_solve(int n, const orc_cplx *A, const orc_cplx *b, orc_cplx *x);
void orc_r_grot(double a, double b, double *c, double *s, double *r);
void orc_c_grot(orc_cplx a, orc_cplx b, orc_cplx *c, orc_cplx *s, orc_cplx *r);

/* ---- QMRIDR, MSIDRS, MSQMRIDR ---- */
int orc_r_qmridr(const orc_matrix *A, int n, const double *b, int s, const orc_idrs_opts *o, double *x,
                 orc_stats *st);
int orc_c_qmridr(const orc_matrix *A, int n, const orc_cplx *b, int s, const orc_idrs_opts *o, orc_cplx *x,
                 orc_stats *st);
int orc_r_msidrs(const orc_matrix *A, int n, const double *b, int n_sigma, const double *sigma, int s,
                 const orc_idrs_opts *o, double *x, orc_stats *st);
int orc_c_msidrs(const orc_matrix *A, int n, const orc_cplx *b, int n_sigma, const orc_cplx *sigma, int s,
                 const orc_idrs_opts *o, orc_cplx *x, orc_stats *st);
int orc_r_msqmridr(const orc_matrix *A, int n, const double *b, int n_sigma, const double *sigma, int s,
                   const orc_idrs_opts *o, double *x, orc_stats *st);
int orc_c_msqmridr(const orc_matrix *A, int n, const orc_cplx *b, int n_sigma, const orc_cplx *sigma, int s,
                   const orc_idrs_opts *o, orc_cplx *x, orc_stats *st);

#endif
