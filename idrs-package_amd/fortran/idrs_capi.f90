!
! iso_c_binding interface to libidrs_hip.so (declarations in
! include/idrs_capi.h).  Used by the drop-in modules user_module,
! matrix_module and idrs_module.
!
module idrs_capi

   use iso_c_binding

   implicit none

   integer(c_int), parameter :: IDRS_OK = 0

! Optional arguments of the solvers (idrs_opts_t)
   type, bind(C) :: idrs_opts_t
      integer(c_int32_t) :: has_M1 = 0
      integer(c_int32_t) :: has_tolerance = 0
      integer(c_int32_t) :: has_maximum_iterations = 0
      integer(c_int32_t) :: has_variant = 0
      type(c_ptr)        :: M1 = c_null_ptr
      real(c_double)     :: tolerance = 0._c_double
      integer(c_int32_t) :: maximum_iterations = 0
      integer(c_int32_t) :: variant = 0
      type(c_ptr)        :: x0 = c_null_ptr
      type(c_ptr)        :: U0 = c_null_ptr
      type(c_ptr)        :: omega = c_null_ptr
      integer(c_int32_t) :: n_omega = 0
      integer(c_int32_t) :: resvec_len = 0
      type(c_ptr)        :: resvec = c_null_ptr
      type(c_ptr)        :: PT_raw = c_null_ptr
      integer(c_int64_t) :: p_seed = 0
      integer(c_int32_t) :: device_pointers = 0
      integer(c_int32_t) :: time_spmv = 0
      integer(c_int32_t) :: graph_mode = 0
      integer(c_int32_t) :: H_nritz = 0
      type(c_ptr)        :: H = c_null_ptr
   end type idrs_opts_t

! Optional arguments of QMRIDR, MSIDRS and MSQMRIDR (idrs_krylov_opts_t)
   type, bind(C) :: idrs_krylov_opts_t
      integer(c_int32_t) :: has_M1 = 0
      integer(c_int32_t) :: has_tolerance = 0
      integer(c_int32_t) :: has_maximum_iterations = 0
      integer(c_int32_t) :: has_variant = 0
      type(c_ptr)        :: M1 = c_null_ptr
      real(c_double)     :: tolerance = 0._c_double
      integer(c_int32_t) :: maximum_iterations = 0
      integer(c_int32_t) :: variant = 0
      integer(c_int32_t) :: has_inner_s = 0
      integer(c_int32_t) :: inner_s = 0
      integer(c_int32_t) :: has_inner_tolerance = 0
      integer(c_int32_t) :: has_inner_iterations = 0
      real(c_double)     :: inner_tolerance = 0._c_double
      integer(c_int32_t) :: inner_iterations = 0
      integer(c_int32_t) :: n_omega = 0
      type(c_ptr)        :: x0 = c_null_ptr
      type(c_ptr)        :: omega = c_null_ptr
      type(c_ptr)        :: resvec = c_null_ptr
      integer(c_int32_t) :: resvec_len = 0
      integer(c_int32_t) :: device_pointers = 0
      type(c_ptr)        :: PT_raw = c_null_ptr
      integer(c_int64_t) :: p_seed = 0
      type(c_ptr)        :: colfac = c_null_ptr
   end type idrs_krylov_opts_t

   type, bind(C) :: idrs_stats_t
      integer(c_int32_t) :: flag = -1
      integer(c_int32_t) :: iterations = 0
      real(c_double)     :: relres = 1._c_double
      real(c_double)     :: solve_ms = 0._c_double
      real(c_double)     :: spmv_ms = 0._c_double
      integer(c_int32_t) :: spmv_launches = 0
      integer(c_int32_t) :: reserved = 0
   end type idrs_stats_t

   interface

      function idrs_last_error() bind(C, name="idrs_last_error") result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function

      integer(c_int) function idrs_crs_format_f64(nrows, ncols, row_ptr, col_ind, rval, A) &
            bind(C, name="idrs_crs_format_f64")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nrows, ncols
         integer(c_int), intent(in) :: row_ptr(*), col_ind(*)
         real(c_double), intent(in), optional :: rval(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_crs_format_c128(nrows, ncols, row_ptr, col_ind, cval, A) &
            bind(C, name="idrs_crs_format_c128")
         import :: c_int, c_double_complex, c_ptr
         integer(c_int), value :: nrows, ncols
         integer(c_int), intent(in) :: row_ptr(*), col_ind(*)
         complex(c_double_complex), intent(in) :: cval(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_mm_info(path, rep, field, symm, nrows, ncols, entries) &
            bind(C, name="idrs_mm_info")
         import :: c_int, c_char, c_int64_t
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), intent(out) :: rep, field, symm, nrows, ncols
         integer(c_int64_t), intent(out) :: entries
      end function

      integer(c_int) function idrs_mm_matrix(path, format, A) bind(C, name="idrs_mm_matrix")
         import :: c_int, c_char, c_ptr
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: format
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_user_matrix_info(A, nrows, ncols, nnz, format) &
            bind(C, name="idrs_user_matrix_info")
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: A
         integer(c_int), intent(out) :: nrows, ncols, format
         integer(c_int64_t), intent(out) :: nnz
      end function

      integer(c_int) function idrs_mm_dense(path, vals) bind(C, name="idrs_mm_dense")
         import :: c_int, c_char, c_double
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), intent(out) :: vals(*)
      end function

      integer(c_int) function idrs_coo_format_f64(nrows, ncols, nnz, indx, jndx, rval, A) &
            bind(C, name="idrs_coo_format_f64")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nrows, ncols, nnz
         integer(c_int), intent(in) :: indx(*), jndx(*)
         real(c_double), intent(in), optional :: rval(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_coo_format_c128(nrows, ncols, nnz, indx, jndx, cval, A) &
            bind(C, name="idrs_coo_format_c128")
         import :: c_int, c_double_complex, c_ptr
         integer(c_int), value :: nrows, ncols, nnz
         integer(c_int), intent(in) :: indx(*), jndx(*)
         complex(c_double_complex), intent(in) :: cval(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_cds_format_f64(nrows, ncols, half_band, rband, A) &
            bind(C, name="idrs_cds_format_f64")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nrows, ncols, half_band
         real(c_double), intent(in), optional :: rband(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_cds_format_c128(nrows, ncols, half_band, cband, A) &
            bind(C, name="idrs_cds_format_c128")
         import :: c_int, c_double_complex, c_ptr
         integer(c_int), value :: nrows, ncols, half_band
         complex(c_double_complex), intent(in) :: cband(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_dia_format_f64(nrows, ncols, ndiag, offsets, vals, A) &
            bind(C, name="idrs_dia_format_f64")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nrows, ncols, ndiag
         integer(c_int), intent(in) :: offsets(*)
         real(c_double), intent(in), optional :: vals(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_dia_format_c128(nrows, ncols, ndiag, offsets, vals, A) &
            bind(C, name="idrs_dia_format_c128")
         import :: c_int, c_double_complex, c_ptr
         integer(c_int), value :: nrows, ncols, ndiag
         integer(c_int), intent(in) :: offsets(*)
         complex(c_double_complex), intent(in) :: vals(*)
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_user_matrix_destroy(A) bind(C, name="idrs_user_matrix_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: A
      end function

      integer(c_int) function idrs_user_mv_f64(K, v, w) bind(C, name="idrs_user_mv_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: K
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_user_mv_c128(K, v, w) bind(C, name="idrs_user_mv_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: K
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_real_diagonal(K, D) bind(C, name="idrs_real_diagonal")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: K
         real(c_double), intent(out) :: D(*)
      end function

      integer(c_int) function idrs_complex_diagonal(K, D) bind(C, name="idrs_complex_diagonal")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: K
         complex(c_double_complex), intent(out) :: D(*)
      end function

      integer(c_int) function idrs_matrix_create(K, A) bind(C, name="idrs_matrix_create")
         import :: c_int, c_ptr
         type(c_ptr), value :: K
         type(c_ptr), intent(out) :: A
      end function

      integer(c_int) function idrs_matrix_destroy(A) bind(C, name="idrs_matrix_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: A
      end function

      integer(c_int) function idrs_matrix_set_precon_f64(A, which, shift) bind(C, name="idrs_matrix_set_precon_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: which
         real(c_double), intent(in), optional :: shift
      end function

      integer(c_int) function idrs_matrix_set_precon_c128(A, which, shift) bind(C, name="idrs_matrix_set_precon_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: which
         complex(c_double_complex), intent(in), optional :: shift
      end function

      integer(c_int) function idrs_matrix_set_seed_f64(A, seed) bind(C, name="idrs_matrix_set_seed_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), value :: seed
      end function

      integer(c_int) function idrs_matrix_set_seed_c128(A, seed) bind(C, name="idrs_matrix_set_seed_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: seed
      end function

      integer(c_int) function idrs_matrix_set_alpha(A, degree, alpha) bind(C, name="idrs_matrix_set_alpha")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: degree
         complex(c_double_complex), intent(in) :: alpha(*)
      end function

      integer(c_int) function idrs_chebyshev_coeffs(degree, foci, is_complex, alpha) &
            bind(C, name="idrs_chebyshev_coeffs")
         import :: c_int, c_double_complex
         integer(c_int), value :: degree, is_complex
         complex(c_double_complex), intent(in) :: foci(*)
         complex(c_double_complex), intent(out) :: alpha(*)
      end function

      integer(c_int) function idrs_neumann_coeffs(degree, center, is_complex, alpha) &
            bind(C, name="idrs_neumann_coeffs")
         import :: c_int, c_double_complex
         integer(c_int), value :: degree, is_complex
         complex(c_double_complex), intent(in) :: center
         complex(c_double_complex), intent(out) :: alpha(*)
      end function

      integer(c_int) function idrs_compute_alpha_coeffs(degree, H, ldh, ncol, is_complex, alpha) &
            bind(C, name="idrs_compute_alpha_coeffs")
         import :: c_int, c_double_complex
         integer(c_int), value :: degree, ldh, ncol, is_complex
         complex(c_double_complex), intent(in) :: H(*)
         complex(c_double_complex), intent(out) :: alpha(*)
      end function

      integer(c_int) function idrs_matvec_f64(A, v, w) bind(C, name="idrs_matvec_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_matvec_c128(A, v, w) bind(C, name="idrs_matvec_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_precon_f64(A, v, w) bind(C, name="idrs_precon_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_precon_c128(A, v, w) bind(C, name="idrs_precon_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_ms_scaleback_f64(A, n_sigma, degree, alpha_s, Y, X, device_pointers) &
            bind(C, name="idrs_ms_scaleback_f64")
         import :: c_int, c_double, c_ptr, c_double_complex
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma, degree, device_pointers
         complex(c_double_complex), intent(in) :: alpha_s(*)
         real(c_double), intent(in) :: Y(*)
         real(c_double), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_ms_scaleback_c128(A, n_sigma, degree, alpha_s, Y, X, device_pointers) &
            bind(C, name="idrs_ms_scaleback_c128")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma, degree, device_pointers
         complex(c_double_complex), intent(in) :: alpha_s(*)
         complex(c_double_complex), intent(in) :: Y(*)
         complex(c_double_complex), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_scaleback_f64(A, y, x) bind(C, name="idrs_scaleback_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: y(*)
         real(c_double), intent(out) :: x(*)
      end function

      integer(c_int) function idrs_scaleback_c128(A, y, x) bind(C, name="idrs_scaleback_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: y(*)
         complex(c_double_complex), intent(out) :: x(*)
      end function

! generalized operators: mass, damping, block systems, deflation
      integer(c_int) function idrs_matrix_set_mass(A, M) bind(C, name="idrs_matrix_set_mass")
         import :: c_int, c_ptr
         type(c_ptr), value :: A, M
      end function

      integer(c_int) function idrs_matrix_set_damping(A, C) bind(C, name="idrs_matrix_set_damping")
         import :: c_int, c_ptr
         type(c_ptr), value :: A, C
      end function

      integer(c_int) function idrs_matrix_set_flags(A, block_system, multishift) bind(C, name="idrs_matrix_set_flags")
         import :: c_int, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: block_system, multishift
      end function

      integer(c_int) function idrs_matrix_rows(A, rows) bind(C, name="idrs_matrix_rows")
         import :: c_int, c_ptr
         type(c_ptr), value :: A
         integer(c_int), intent(out) :: rows
      end function

      integer(c_int) function idrs_matrix_set_precon_diag(A, PK, PM, PC, shift, is_complex) &
            bind(C, name="idrs_matrix_set_precon_diag")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: A, PK, PM, PC
         real(c_double), intent(in), optional :: shift(*)
         integer(c_int), value :: is_complex
      end function

      integer(c_int) function idrs_matrix_set_precon_lu(A, which, PK, PM, PC, shift, is_complex) &
            bind(C, name="idrs_matrix_set_precon_lu")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: A, PK, PM, PC
         integer(c_int), value :: which, is_complex
         real(c_double), intent(in), optional :: shift(*)
      end function

      integer(c_int) function idrs_matrix_share_lu(A, B) bind(C, name="idrs_matrix_share_lu")
         import :: c_int, c_ptr
         type(c_ptr), value :: A, B
      end function

      integer(c_int) function idrs_matrix_set_shift_invert(A, shift_invert, shift, is_complex) &
            bind(C, name="idrs_matrix_set_shift_invert")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: A
         integer(c_int), value :: shift_invert, is_complex
         real(c_double), intent(in), optional :: shift(*)
      end function

      integer(c_int) function idrs_make_deflation_f64(A, nd, V) bind(C, name="idrs_make_deflation_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: nd
         real(c_double), intent(in) :: V(*)
      end function

      integer(c_int) function idrs_make_deflation_c128(A, nd, V) bind(C, name="idrs_make_deflation_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         integer(c_int), value :: nd
         complex(c_double_complex), intent(in) :: V(*)
      end function

      integer(c_int) function idrs_pmul_f64(A, v, w) bind(C, name="idrs_pmul_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_pmul_c128(A, v, w) bind(C, name="idrs_pmul_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_system_mv_f64(A, v, w) bind(C, name="idrs_system_mv_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_system_mv_c128(A, v, w) bind(C, name="idrs_system_mv_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: w(*)
      end function

      integer(c_int) function idrs_add_deflation_f64(A, b, y, x) bind(C, name="idrs_add_deflation_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: b(*), y(*)
         real(c_double), intent(out) :: x(*)
      end function

      integer(c_int) function idrs_add_deflation_c128(A, b, y, x) bind(C, name="idrs_add_deflation_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: b(*), y(*)
         complex(c_double_complex), intent(out) :: x(*)
      end function

      integer(c_int) function idrs_scaleback_b_f64(A, b, y, x) bind(C, name="idrs_scaleback_b_f64")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: A
         real(c_double), intent(in) :: b(*), y(*)
         real(c_double), intent(out) :: x(*)
      end function

      integer(c_int) function idrs_scaleback_b_c128(A, b, y, x) bind(C, name="idrs_scaleback_b_c128")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: A
         complex(c_double_complex), intent(in) :: b(*), y(*)
         complex(c_double_complex), intent(out) :: x(*)
      end function

      integer(c_int) function idrs_ms_scaleback_b_f64(A, n_sigma, sigma, degree, alpha_s, b, Y, X) &
            bind(C, name="idrs_ms_scaleback_b_f64")
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma, degree
         real(c_double), intent(in) :: sigma(*), b(*), Y(*)
         complex(c_double_complex), intent(in) :: alpha_s(*)
         real(c_double), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_ms_scaleback_b_c128(A, n_sigma, sigma, degree, alpha_s, b, Y, X) &
            bind(C, name="idrs_ms_scaleback_b_c128")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma, degree
         complex(c_double_complex), intent(in) :: sigma(*), b(*), Y(*), alpha_s(*)
         complex(c_double_complex), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_ms_add_deflation_f64(A, n_sigma, sigma, b, Y, X) &
            bind(C, name="idrs_ms_add_deflation_f64")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma
         real(c_double), intent(in) :: sigma(*), b(*), Y(*)
         real(c_double), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_ms_add_deflation_c128(A, n_sigma, sigma, b, Y, X) &
            bind(C, name="idrs_ms_add_deflation_c128")
         import :: c_int, c_ptr, c_double_complex
         type(c_ptr), value :: A
         integer(c_int), value :: n_sigma
         complex(c_double_complex), intent(in) :: sigma(*), b(*), Y(*)
         complex(c_double_complex), intent(out) :: X(*)
      end function

      integer(c_int) function idrs_idrs_f64(A, n, b, s, opts, x, stats) bind(C, name="idrs_idrs_f64")
         import :: c_int, c_double, c_ptr, idrs_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s
         real(c_double), intent(in) :: b(*)
         type(idrs_opts_t), intent(in) :: opts
         real(c_double), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_idrs_c128(A, n, b, s, opts, x, stats) bind(C, name="idrs_idrs_c128")
         import :: c_int, c_double_complex, c_ptr, idrs_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s
         complex(c_double_complex), intent(in) :: b(*)
         type(idrs_opts_t), intent(in) :: opts
         complex(c_double_complex), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_idrs_ritz_f64(A, n, b, s, opts, x, stats, omega, n_omega, U0, dim_u0, &
                                                 H, h_rows, h_cols, nritz, converged) &
            bind(C, name="idrs_idrs_ritz_f64")
         import :: c_int, c_double, c_ptr, idrs_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s, n_omega, dim_u0, h_rows, h_cols
         real(c_double), intent(in) :: b(*)
         type(idrs_opts_t), intent(in) :: opts
         real(c_double), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
         type(c_ptr), value :: omega, U0, H
         integer(c_int), intent(out) :: nritz, converged
      end function

      integer(c_int) function idrs_idrs_ritz_c128(A, n, b, s, opts, x, stats, omega, n_omega, U0, dim_u0, &
                                                  H, h_rows, h_cols, nritz, converged) &
            bind(C, name="idrs_idrs_ritz_c128")
         import :: c_int, c_double_complex, c_ptr, idrs_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s, n_omega, dim_u0, h_rows, h_cols
         complex(c_double_complex), intent(in) :: b(*)
         type(idrs_opts_t), intent(in) :: opts
         complex(c_double_complex), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
         type(c_ptr), value :: omega, U0, H
         integer(c_int), intent(out) :: nritz, converged
      end function

! Spectral routines on small dense matrices (complex values, pairs)
      integer(c_int) function idrs_qr_eig(n, H, ldh, is_complex, eig, U, converged) bind(C, name="idrs_qr_eig")
         import :: c_int, c_double_complex, c_ptr
         integer(c_int), value :: n, ldh, is_complex
         complex(c_double_complex), intent(in) :: H(*)
         complex(c_double_complex), intent(out) :: eig(*)
         type(c_ptr), value :: U
         integer(c_int), intent(out) :: converged
      end function

      integer(c_int) function idrs_filter_ritzval(n, ritzval, outlier) bind(C, name="idrs_filter_ritzval")
         import :: c_int, c_double_complex
         integer(c_int), value :: n
         complex(c_double_complex), intent(in) :: ritzval(*)
         integer(c_int), intent(out) :: outlier(*)
      end function

      integer(c_int) function idrs_analyse_spectrum(n, H, ldh, is_complex, foci, center, shift) &
            bind(C, name="idrs_analyse_spectrum")
         import :: c_int, c_double_complex
         integer(c_int), value :: n, ldh, is_complex
         complex(c_double_complex), intent(in) :: H(*)
         complex(c_double_complex), intent(out) :: foci(2), center, shift
      end function

      integer(c_int) function idrs_si_power(n, H, ldh, is_complex, s, shift, seed, offset, schurvec) &
            bind(C, name="idrs_si_power")
         import :: c_int, c_int64_t, c_double_complex
         integer(c_int), value :: n, ldh, is_complex, s
         complex(c_double_complex), intent(in) :: H(*), shift
         integer(c_int64_t), value :: seed, offset
         complex(c_double_complex), intent(out) :: schurvec(*)
      end function

      integer(c_int) function idrs_qmridr_f64(A, n, b, s, opts, x, stats) bind(C, name="idrs_qmridr_f64")
         import :: c_int, c_double, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s
         real(c_double), intent(in) :: b(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         real(c_double), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_qmridr_c128(A, n, b, s, opts, x, stats) bind(C, name="idrs_qmridr_c128")
         import :: c_int, c_double_complex, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, s
         complex(c_double_complex), intent(in) :: b(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         complex(c_double_complex), intent(out) :: x(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_msidrs_f64(A, n, b, n_sigma, sigma, s, opts, X, relres, stats) &
            bind(C, name="idrs_msidrs_f64")
         import :: c_int, c_double, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, n_sigma, s
         real(c_double), intent(in) :: b(*), sigma(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         real(c_double), intent(out) :: X(*)
         real(c_double), intent(out) :: relres(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_msidrs_c128(A, n, b, n_sigma, sigma, s, opts, X, relres, stats) &
            bind(C, name="idrs_msidrs_c128")
         import :: c_int, c_double_complex, c_double, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, n_sigma, s
         complex(c_double_complex), intent(in) :: b(*), sigma(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         complex(c_double_complex), intent(out) :: X(*)
         real(c_double), intent(out) :: relres(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_msqmridr_f64(A, n, b, n_sigma, sigma, s, opts, X, relres, stats) &
            bind(C, name="idrs_msqmridr_f64")
         import :: c_int, c_double, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, n_sigma, s
         real(c_double), intent(in) :: b(*), sigma(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         real(c_double), intent(out) :: X(*)
         real(c_double), intent(out) :: relres(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

      integer(c_int) function idrs_msqmridr_c128(A, n, b, n_sigma, sigma, s, opts, X, relres, stats) &
            bind(C, name="idrs_msqmridr_c128")
         import :: c_int, c_double_complex, c_double, c_ptr, idrs_krylov_opts_t, idrs_stats_t
         type(c_ptr), value :: A
         integer(c_int), value :: n, n_sigma, s
         complex(c_double_complex), intent(in) :: b(*), sigma(*)
         type(idrs_krylov_opts_t), intent(in) :: opts
         complex(c_double_complex), intent(out) :: X(*)
         real(c_double), intent(out) :: relres(*)
         type(idrs_stats_t), intent(out) :: stats
      end function

   end interface

contains

! Stops with the library's message when a call failed (the reference's
! "error stop" behaviour for illegal input).
   subroutine idrs_check(status, where)
      integer(c_int), intent(in) :: status
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      character(len=:), allocatable :: text
      integer :: i
      type(c_ptr) :: p
      if ( status == IDRS_OK ) return
      p = idrs_last_error()
      text = ''
      if ( c_associated(p) ) then
         call c_f_pointer(p, msg, [1024])
         do i = 1, 1024
            if ( msg(i) == c_null_char ) exit
            text = text // msg(i)
         end do
      end if
      write(*,'(a,a,a,i0,a,a)') 'IDRS library error in ', where, ' (status ', status, '): ', text
      error stop 'IDRS library error'
   end subroutine idrs_check

end module idrs_capi
