!
! user_module: drop-in for the reference's MODULES/PP_MATRIX/user_module.f90
! public interface on the hot path.  type(user_matrix) keeps the reference's
! scalar components and holds a handle to device storage in libidrs_hip.so;
! the products run in the library's HIP kernels.
!
!   crs_format / coo_format / cds_format    user_module.f90:829-958
!   operator(*)  (ruser_mv / cuser_mv)       user_module.f90:197-241
!   operator(/)  (ruser_precon / cuser_precon: diagonal, LU)  :243-283
!   real_diagonal / complex_diagonal         user_module.f90:960-1161
!   REAL_PRECON / COMPLEX_PRECON(which, K, M, C, shift)
!                 1 diagonal, 2 dense LU, 3 band LU  user_module.f90:1327-1602
!
! Storage formats: matrix_format 2 CRS, 3 COO, 4 CDS, 5 DIA (offset list,
! an extension for wide sparse bands); 1 dense holds host data only (dense
! products stop with a message).  The dense and band LU preconditioners
! (which_precon 2, 3) keep their factors on the device (the reference's
! P%rmat / P%rband host copies are not filled); the dense one takes CRS or
! COO matrices.  Device storage is released by user_matrix_free (the reference's
! allocatable components are freed with the object; these handles are not).
!
module user_module

   use precision_module
   use idrs_capi
   use iso_c_binding

   implicit none

   private
   public :: user_matrix, operator(*), operator(/), crs_format, coo_format, cds_format, dia_format, &
             dense_format, real_diagonal, complex_diagonal, REAL_PRECON, COMPLEX_PRECON, user_matrix_free, &
             user_matrix_sizes

   type user_matrix
      integer                        :: matrix_format = 0
      integer                        :: which_precon = 0
      integer                        :: nrows = 0           ! Number of rows
      integer                        :: ncols = 0           ! Number columns
      integer                        :: nnz = 0             ! Number of stored nonzeros
      integer                        :: half_band = 0       ! CDS half bandwidth
      logical                        :: is_complex = .false.
! Diagonal preconditioner (which_precon = 1): band(:,0) as in the reference
      real(kind=rp), allocatable     :: rband(:,:)
      complex(kind=cp), allocatable  :: cband(:,:)
      logical                        :: has_shift = .false.
      real(kind=rp)                  :: rshift = 0._rp
      complex(kind=cp)               :: cshift = (0._cp,0._cp)
! Dense storage (matrix_format 1): host data such as a block of right-hand
! sides read from a file; dense products are not offered
      real(kind=rp), allocatable     :: rmat(:,:)
      complex(kind=cp), allocatable  :: cmat(:,:)
      integer, allocatable           :: imat(:,:)
! Device objects: h is the stored matrix, hp the diagonal preconditioner
! operator built from h (which_precon = 1); hm, hc the mass and damping
! matrices whose diagonals enter a shifted diagonal preconditioner
      type(c_ptr)                    :: h = c_null_ptr
      type(c_ptr)                    :: hp = c_null_ptr
      type(c_ptr)                    :: hm = c_null_ptr
      type(c_ptr)                    :: hc = c_null_ptr
   end type user_matrix

   INTERFACE OPERATOR(*)
      module procedure ruser_mv, cuser_mv
   END INTERFACE

   INTERFACE OPERATOR(/)
      module procedure ruser_precon, cuser_precon
   END INTERFACE

contains

   function ruser_mv( A, v ) result(w)
      type(user_matrix), intent(in)     :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(A%nrows)
      if ( A%matrix_format == 1 ) stop 'dense products are not offered by the GPU library, use crs_format'
      if ( .not. c_associated(A%h) ) stop 'No matrix format specified.'
      call idrs_check( idrs_user_mv_f64(A%h, v, w), 'user_mv' )
   end function ruser_mv

   function cuser_mv( A, v ) result(w)
      type(user_matrix), intent(in)     :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(A%nrows)
      if ( A%matrix_format == 1 ) stop 'dense products are not offered by the GPU library, use crs_format'
      if ( .not. c_associated(A%h) ) stop 'Illegal value for matrix format'
      call idrs_check( idrs_user_mv_c128(A%h, v, w), 'user_mv' )
   end function cuser_mv

   function ruser_precon( v, A ) result(w)
      type(user_matrix), intent(in)     :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      if ( A%which_precon >= 1 ) then
         call idrs_check( idrs_precon_f64(A%hp, v, w), 'user_precon' )
      else
         w = v
      end if
   end function ruser_precon

   function cuser_precon( v, A ) result(w)
      type(user_matrix), intent(in)     :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      if ( A%which_precon >= 1 ) then
         call idrs_check( idrs_precon_c128(A%hp, v, w), 'user_precon' )
      else
         w = v
      end if
   end function cuser_precon

   function crs_format( nrows, ncols, row_ptr, col_ind, rval, ival, cval ) result(A)
      type(user_matrix)                      :: A
      integer                                :: nrows, ncols
      integer                                :: row_ptr(:), col_ind(:)
      real(kind=rp), optional, intent(in)    :: rval(:)
      integer, optional, intent(in)          :: ival(:)
      complex(kind=cp), optional, intent(in) :: cval(:)
      real(kind=rp), allocatable             :: tmp(:)

      A%matrix_format = 2
      A%nrows = nrows
      A%ncols = ncols
      A%nnz = row_ptr(nrows+1) - 1
      if ( present(cval) ) then
         A%is_complex = .true.
         call idrs_check( idrs_crs_format_c128(nrows, ncols, row_ptr, col_ind, cval, A%h), 'crs_format' )
      elseif ( present(rval) ) then
         call idrs_check( idrs_crs_format_f64(nrows, ncols, row_ptr, col_ind, rval, A%h), 'crs_format' )
      elseif ( present(ival) ) then
! integer values are stored as real values (user_module.f90:860)
         tmp = real(ival, kind=rp)
         call idrs_check( idrs_crs_format_f64(nrows, ncols, row_ptr, col_ind, tmp, A%h), 'crs_format' )
      else
         call idrs_check( idrs_crs_format_f64(nrows, ncols, row_ptr, col_ind, A=A%h), 'crs_format' )
      end if
   end function crs_format

   function coo_format( nrows, ncols, nnz, indx, jndx, rval, ival, cval ) result(A)
      type(user_matrix)                      :: A
      integer                                :: nrows, ncols, nnz
      integer                                :: indx(:), jndx(:)
      real(kind=rp), optional, intent(in)    :: rval(:)
      integer, optional, intent(in)          :: ival(:)
      complex(kind=cp), optional, intent(in) :: cval(:)
      real(kind=rp), allocatable             :: tmp(:)

      A%matrix_format = 3
      A%nrows = nrows
      A%ncols = ncols
      A%nnz   = nnz
      if ( present(cval) ) then
         A%is_complex = .true.
         call idrs_check( idrs_coo_format_c128(nrows, ncols, nnz, indx, jndx, cval, A%h), 'coo_format' )
      elseif ( present(rval) ) then
         call idrs_check( idrs_coo_format_f64(nrows, ncols, nnz, indx, jndx, rval, A%h), 'coo_format' )
      elseif ( present(ival) ) then
         tmp = real(ival(1:nnz), kind=rp)
         call idrs_check( idrs_coo_format_f64(nrows, ncols, nnz, indx, jndx, tmp, A%h), 'coo_format' )
      else
         call idrs_check( idrs_coo_format_f64(nrows, ncols, nnz, indx, jndx, A=A%h), 'coo_format' )
      end if
   end function coo_format

   function cds_format( nrows, ncols, half_band, rband, iband, cband ) result(A)
      type(user_matrix)                      :: A
      integer                                :: nrows, ncols, half_band
      real(kind=rp), optional                :: rband(nrows,-half_band:half_band)
      integer, optional                      :: iband(nrows,-half_band:half_band)
      complex(kind=cp), optional             :: cband(nrows,-half_band:half_band)
      real(kind=rp), allocatable             :: tmp(:,:)

      A%matrix_format = 4
      A%nrows = nrows
      A%ncols = ncols
      A%half_band = half_band
      A%nnz = nrows*(2*half_band+1)
      if ( present(cband) ) then
         A%is_complex = .true.
         call idrs_check( idrs_cds_format_c128(nrows, ncols, half_band, cband, A%h), 'cds_format' )
      elseif ( present(rband) ) then
         call idrs_check( idrs_cds_format_f64(nrows, ncols, half_band, rband, A%h), 'cds_format' )
      elseif ( present(iband) ) then
         tmp = real(iband, kind=rp)
         call idrs_check( idrs_cds_format_f64(nrows, ncols, half_band, tmp, A%h), 'cds_format' )
      else
         call idrs_check( idrs_cds_format_f64(nrows, ncols, half_band, A=A%h), 'cds_format' )
      end if
   end function cds_format

! Offset-list diagonal storage: vals(nrows, ndiag) for strictly ascending
! offsets; entry (i, d) multiplies v(i + offsets(d)).
   function dia_format( nrows, ncols, offsets, rvals, cvals ) result(A)
      type(user_matrix)                      :: A
      integer                                :: nrows, ncols
      integer                                :: offsets(:)
      real(kind=rp), optional                :: rvals(:,:)
      complex(kind=cp), optional             :: cvals(:,:)

      A%matrix_format = 5
      A%nrows = nrows
      A%ncols = ncols
      A%nnz = nrows*size(offsets)
      if ( present(cvals) ) then
         A%is_complex = .true.
         call idrs_check( idrs_dia_format_c128(nrows, ncols, size(offsets), offsets, cvals, A%h), 'dia_format' )
      elseif ( present(rvals) ) then
         call idrs_check( idrs_dia_format_f64(nrows, ncols, size(offsets), offsets, rvals, A%h), 'dia_format' )
      else
         call idrs_check( idrs_dia_format_f64(nrows, ncols, size(offsets), offsets, A=A%h), 'dia_format' )
      end if
   end function dia_format

! dense_format (user_module.f90:900-930): the dense matrix as host data
! (its values, e.g. the right-hand sides of a Matrix Market array file);
! the products of a dense matrix are not offered by the GPU library and stop
! with a message
   function dense_format( rmat, imat, cmat ) result(A)
      type(user_matrix)                      :: A
      real(kind=rp), optional                :: rmat(:,:)
      integer, optional                      :: imat(:,:)
      complex(kind=cp), optional             :: cmat(:,:)
      A%matrix_format = 1
      if ( present(rmat) ) then
         A%nrows = size(rmat,1)
         A%ncols = size(rmat,2)
         A%rmat = rmat
      elseif ( present(cmat) ) then
         A%nrows = size(cmat,1)
         A%ncols = size(cmat,2)
         A%is_complex = .true.
         A%cmat = cmat
      elseif ( present(imat) ) then
         A%nrows = size(imat,1)
         A%ncols = size(imat,2)
         A%imat = imat
      end if
      A%nnz = A%nrows*A%ncols
   end function dense_format

! nrows, ncols and stored entries of a user matrix from its device object
   subroutine user_matrix_sizes( A )
      type(user_matrix), intent(inout)      :: A
      integer(c_int)                        :: nr, nc, fmt
      integer(c_int64_t)                    :: nz
      call idrs_check( idrs_user_matrix_info(A%h, nr, nc, nz, fmt), 'user_matrix_info' )
      A%nrows = nr
      A%ncols = nc
      A%nnz = int(nz)
   end subroutine user_matrix_sizes

   function real_diagonal( A ) result(D)
      type(user_matrix), intent(in)         :: A
      real(kind=rp)                         :: D(A%nrows)
      call idrs_check( idrs_real_diagonal(A%h, D), 'real_diagonal' )
   end function real_diagonal

   function complex_diagonal( A ) result(D)
      type(user_matrix), intent(in)         :: A
      complex(kind=cp)                      :: D(A%nrows)
      call idrs_check( idrs_complex_diagonal(A%h, D), 'complex_diagonal' )
   end function complex_diagonal

   function REAL_PRECON( which_precon, K, M, C, shift ) result(P)
!
! which_precon = 0 : no preconditioner
! which_precon = 1 : diagonal preconditioner (user_module.f90:1352-1381):
!    D = diag(K) + shift diag(C) + shift^2 diag(M)   (M and C)
!        diag(K) + shift diag(C) + shift^2           (C)
!        diag(K) - shift diag(M)                     (M)
!        diag(K) - shift                              (neither)
!    and diag(K) without a shift, whatever M and C
!
      integer, intent(in)                              :: which_precon
      type(user_matrix), intent(in)                    :: K
      type(user_matrix), intent(in), optional          :: M
      type(user_matrix), intent(in), optional          :: C
      type(user_matrix)                                :: P
      real(kind=rp), optional                          :: shift
      real(c_double)                                   :: sh(1)

      P%matrix_format = 0
      P%which_precon = which_precon
      P%nrows = K%nrows
      P%ncols = K%ncols
      if ( which_precon == 1 ) then
         P%matrix_format = 4
         P%h = K%h
         call idrs_check( idrs_matrix_create(K%h, P%hp), 'REAL_PRECON' )
         allocate( P%rband(P%nrows,0:0) )
         P%rband(:,0) = real_diagonal( K )
         if ( present(shift) ) then
            P%has_shift = .true.
            P%rshift = shift
            if ( present(M) ) P%hm = M%h
            if ( present(C) ) P%hc = C%h
            if ( present(M) .or. present(C) ) then
               sh(1) = shift
               call idrs_check( idrs_matrix_set_precon_diag(P%hp, K%h, P%hm, P%hc, sh, 0), 'REAL_PRECON' )
            else
               call idrs_check( idrs_matrix_set_precon_f64(P%hp, 1, shift), 'REAL_PRECON' )
            end if
! rband(:,0) = the diagonal the preconditioner divides by
            if ( present(M) .and. present(C) ) then
               P%rband(:,0) = P%rband(:,0) + shift*real_diagonal(C) + shift**2*real_diagonal(M)
            elseif ( present(C) ) then
               P%rband(:,0) = P%rband(:,0) + shift*real_diagonal(C) + shift**2
            elseif ( present(M) ) then
               P%rband(:,0) = P%rband(:,0) - shift*real_diagonal(M)
            else
               P%rband(:,0) = P%rband(:,0) - shift
            end if
         else
            call idrs_check( idrs_matrix_set_precon_f64(P%hp, 1), 'REAL_PRECON' )
         end if
      elseif ( which_precon == 2 .or. which_precon == 3 ) then
! dense (2) or band (3) LU of K [+ shift C + shift^2 M | ...] (:1383-1459),
! factored and kept on the device: v/P solves with the factors
         P%matrix_format = merge(1, 4, which_precon == 2)
         P%h = K%h
         if ( present(M) ) P%hm = M%h
         if ( present(C) ) P%hc = C%h
         call idrs_check( idrs_matrix_create(K%h, P%hp), 'REAL_PRECON' )
         if ( present(shift) ) then
            P%has_shift = .true.
            P%rshift = shift
            sh(1) = shift
            call idrs_check( idrs_matrix_set_precon_lu(P%hp, which_precon, K%h, P%hm, P%hc, sh, 0), 'REAL_PRECON' )
         else
            call idrs_check( idrs_matrix_set_precon_lu(P%hp, which_precon, K%h, P%hm, P%hc, is_complex=0), &
                             'REAL_PRECON' )
         end if
      elseif ( which_precon /= 0 ) then
         error stop 'REAL_PRECON: which_precon must be 0, 1, 2 or 3'
      end if
   end function REAL_PRECON

   function COMPLEX_PRECON( which_precon, K, M, C, shift ) result(P)
!
! COMPLEX_PRECON (user_module.f90:1465-1530): the diagonals of REAL_PRECON
! with a complex shift
!
      integer, intent(in)                              :: which_precon
      type(user_matrix), intent(in)                    :: K
      type(user_matrix), intent(in), optional          :: M
      type(user_matrix), intent(in), optional          :: C
      type(user_matrix)                                :: P
      complex(kind=cp), optional                       :: shift
      real(c_double)                                   :: sh(2)

      P%matrix_format = 0
      P%which_precon = which_precon
      P%nrows = K%nrows
      P%ncols = K%ncols
      P%is_complex = .true.
      if ( which_precon == 1 ) then
         P%matrix_format = 4
         P%h = K%h
         call idrs_check( idrs_matrix_create(K%h, P%hp), 'COMPLEX_PRECON' )
         allocate( P%cband(P%nrows,0:0) )
         P%cband(:,0) = complex_diagonal( K )
         if ( present(shift) ) then
            P%has_shift = .true.
            P%cshift = shift
            if ( present(M) ) P%hm = M%h
            if ( present(C) ) P%hc = C%h
            if ( present(M) .or. present(C) ) then
               sh = [ real(shift, kind=rp), aimag(shift) ]
               call idrs_check( idrs_matrix_set_precon_diag(P%hp, K%h, P%hm, P%hc, sh, 1), 'COMPLEX_PRECON' )
            else
               call idrs_check( idrs_matrix_set_precon_c128(P%hp, 1, shift), 'COMPLEX_PRECON' )
            end if
            if ( present(M) .and. present(C) ) then
               P%cband(:,0) = P%cband(:,0) + shift*complex_diagonal(C) + shift**2*complex_diagonal(M)
            elseif ( present(C) ) then
               P%cband(:,0) = P%cband(:,0) + shift*complex_diagonal(C) + shift**2
            elseif ( present(M) ) then
               P%cband(:,0) = P%cband(:,0) - shift*complex_diagonal(M)
            else
               P%cband(:,0) = P%cband(:,0) - shift
            end if
         else
            call idrs_check( idrs_matrix_set_precon_c128(P%hp, 1), 'COMPLEX_PRECON' )
         end if
      elseif ( which_precon == 2 .or. which_precon == 3 ) then
         P%matrix_format = merge(1, 4, which_precon == 2)
         P%h = K%h
         if ( present(M) ) P%hm = M%h
         if ( present(C) ) P%hc = C%h
         call idrs_check( idrs_matrix_create(K%h, P%hp), 'COMPLEX_PRECON' )
         if ( present(shift) ) then
            P%has_shift = .true.
            P%cshift = shift
            sh = [ real(shift, kind=rp), aimag(shift) ]
            call idrs_check( idrs_matrix_set_precon_lu(P%hp, which_precon, K%h, P%hm, P%hc, sh, 1), 'COMPLEX_PRECON' )
         else
            call idrs_check( idrs_matrix_set_precon_lu(P%hp, which_precon, K%h, P%hm, P%hc, is_complex=1), &
                             'COMPLEX_PRECON' )
         end if
      elseif ( which_precon /= 0 ) then
         error stop 'COMPLEX_PRECON: which_precon must be 0, 1, 2 or 3'
      end if
   end function COMPLEX_PRECON

! Releases the device storage behind A (and a preconditioner built from it).
   subroutine user_matrix_free( A )
      type(user_matrix), intent(inout)   :: A
      integer(c_int)                     :: status
      if ( c_associated(A%hp) ) status = idrs_matrix_destroy(A%hp)
      if ( c_associated(A%h) .and. A%which_precon == 0 ) status = idrs_user_matrix_destroy(A%h)
      A%h = c_null_ptr
      A%hp = c_null_ptr
      A%hm = c_null_ptr
      A%hc = c_null_ptr
   end subroutine user_matrix_free

end module user_module
