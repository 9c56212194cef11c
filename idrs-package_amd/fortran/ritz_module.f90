!
! ritz_module: drop-in for the reference's MODULES/IDRS/ritz_module.f90
! (IDRS_RITZ) and the QR_eig / SI_power generics of
! MODULES/IDRS/dense_la_module.f90.  The IDR(s) solve, its Hessenberg
! matrix and the recycle space U0 (compute_U0 and the global QR) run on
! the GPU in libidrs_hip.so; the small dense eigenvalue work runs on the
! host inside the library.
!
!   IDRS_RITZ (RIDRS_RITZ, CIDRS_RITZ)  ritz_module.f90:225-556
!   QR_eig    (RQR_eig, CQR_eig)        dense_la_module.f90:192-403 (no T output)
!   SI_power  (RSI_power, CSI_power)    dense_la_module.f90:455-603
!
! SI_power's start block comes from the library's uniform stream (seed 1,
! after the shadow space's values in IDRS_RITZ), not from the compiler's
! random_number.
!
module ritz_module

   use precision_module
   use matrix_module
   use idrs_module
   use idrs_capi
   use iso_c_binding

   implicit none

   private
   public :: IDRS_RITZ, QR_eig, SI_power

   interface IDRS_RITZ
      module procedure CIDRS_RITZ, RIDRS_RITZ
   end interface

   interface QR_eig
      module procedure CQR_eig, RQR_eig
   end interface

   interface SI_power
      module procedure CSI_power, RSI_power
   end interface

contains

! Shared options of RIDRS_RITZ and CIDRS_RITZ (presence flags as present()).
   subroutine ritz_opts( o, n, s, M, tolerance, maximum_iterations, variant, resvec_len, maxit )
      type(idrs_opts_t), intent(inout)       :: o
      integer, intent(in)                    :: n, s
      type(matrix), optional, intent(in)     :: M
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, intent(in)                    :: resvec_len
      integer, intent(out)                   :: maxit

! v/M divides by M%P's diagonal; a default type(matrix) is the identity
      if ( present(M) ) then
         o%has_M1 = 1
         if ( M%P%which_precon == 1 ) o%M1 = M%P%hp
      end if
      if ( present(tolerance) ) then
         if ( tolerance < 0 ) stop "Illegal value for parameter tolerance"
         o%has_tolerance = 1
         o%tolerance = tolerance
      end if
      maxit = min(2*n,1000)
      if ( present(maximum_iterations) ) then
         o%has_maximum_iterations = 1
         o%maximum_iterations = maximum_iterations
         maxit = maximum_iterations
      end if
      if ( present(variant) ) then
         o%has_variant = 1
         o%variant = variant
      end if
      if ( resvec_len >= 0 ) then
         if ( maxit+1 > resvec_len ) &
            error stop "Length of vector with residual norms too small, should be maxit+1"
      end if
   end subroutine ritz_opts

   function RIDRS_RITZ( A, b, s, &                                     ! required
                        M, tolerance, maximum_iterations, variant, &   ! optional input
                        flag, relres, iterations, &                    ! optional output
                        x0, U0, omega, resvec, H ) result(x)           ! optional arrays, output
      type(matrix), intent(in)                             :: A
      real(kind=rp), intent(in)                            :: b(:)
      integer, intent(in)                                  :: s
      real(kind=rp)                                        :: x(size(b))
      type(matrix), optional, intent(in)                   :: M
      real(kind=rp), optional, intent(in)                  :: tolerance
      integer, optional, intent(in)                        :: maximum_iterations
      integer, optional, intent(in)                        :: variant
      integer, optional, intent(out)                       :: flag
      real(kind=rp), optional, intent(out)                 :: relres
      integer, optional, intent(out)                       :: iterations
      real(kind=rp), optional, intent(in)                  :: x0(:)
      real(kind=rp), optional, allocatable, intent(inout)  :: U0(:,:)
      real(kind=rp), optional, intent(out)                 :: omega(:)
      real(kind=rp), optional                              :: resvec(:)
      real(kind=rp), optional, intent(out)                 :: H(:,:)

      type(idrs_opts_t)                   :: o
      type(idrs_stats_t)                  :: st
      real(kind=rp), allocatable, target  :: bx0(:), brv(:), bom(:), bU0(:,:), bH(:,:)
      real(kind=rp), allocatable          :: bb(:)
      type(c_ptr)                         :: pom, pu, ph
      integer                             :: n, maxit, rvlen, n_omega, dimU0, hr, hc
      integer(c_int)                      :: nritz, converged

      n = size(b)
      rvlen = -1
      if ( present(resvec) ) rvlen = size(resvec)
      call ritz_opts( o, n, s, M, tolerance, maximum_iterations, variant, rvlen, maxit )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      n_omega = 0; dimU0 = 0; hr = 0; hc = 0
      pom = c_null_ptr; pu = c_null_ptr; ph = c_null_ptr
      if ( present(omega) ) then
         n_omega = size(omega)
         allocate( bom(max(n_omega,1)) )
         pom = c_loc(bom)
      end if
      if ( present(U0) ) then
         if ( allocated(U0) ) then
            dimU0 = size(U0,2)
            allocate( bU0(n, max(dimU0,1)) )
            pu = c_loc(bU0)
         end if
      end if
      if ( present(H) ) then
         hr = size(H,1); hc = size(H,2)
         allocate( bH(max(hr,1), max(hc,1)) )
         ph = c_loc(bH)
      end if
      bb = b
      call idrs_check( idrs_idrs_ritz_f64(matrix_handle(A, .false.), n, bb, s, o, x, st, pom, n_omega, pu, dimU0, &
                                          ph, hr, hc, nritz, converged), 'IDRS_RITZ' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( n_omega > 0 ) then
         if ( converged < n_omega ) write(*,'(a,i2,a)') ' Warning, only ',converged, ' Ritzvalues converged. '
         omega = bom(1:n_omega)
      end if
      if ( dimU0 > 0 ) U0 = bU0
      if ( present(H) ) H = bH(1:hr,1:hc)
   end function RIDRS_RITZ

   function CIDRS_RITZ( A, b, s, &                                     ! required
                        M, tolerance, maximum_iterations, variant, &   ! optional input
                        flag, relres, iterations, &                    ! optional output
                        x0, U0, omega, resvec, H ) result(x)           ! optional arrays, output
      type(matrix), intent(in)                               :: A
      complex(kind=cp), intent(in)                           :: b(:)
      integer, intent(in)                                    :: s
      complex(kind=cp)                                       :: x(size(b))
      type(matrix), optional, intent(in)                     :: M
      real(kind=rp), optional, intent(in)                    :: tolerance
      integer, optional, intent(in)                          :: maximum_iterations
      integer, optional, intent(in)                          :: variant
      integer, optional, intent(out)                         :: flag
      real(kind=rp), optional, intent(out)                   :: relres
      integer, optional, intent(out)                         :: iterations
      complex(kind=cp), optional, intent(in)                 :: x0(:)
      complex(kind=cp), optional, allocatable, intent(inout) :: U0(:,:)
      complex(kind=cp), optional, intent(out)                :: omega(:)
      real(kind=rp), optional                                :: resvec(:)
      complex(kind=cp), optional, intent(out)                :: H(:,:)

      type(idrs_opts_t)                     :: o
      type(idrs_stats_t)                    :: st
      complex(kind=cp), allocatable, target :: bx0(:), bom(:), bU0(:,:), bH(:,:)
      real(kind=rp), allocatable, target    :: brv(:)
      complex(kind=cp), allocatable         :: bb(:)
      type(c_ptr)                           :: pom, pu, ph
      integer                               :: n, maxit, rvlen, n_omega, dimU0, hr, hc
      integer(c_int)                        :: nritz, converged

      n = size(b)
      rvlen = -1
      if ( present(resvec) ) rvlen = size(resvec)
      call ritz_opts( o, n, s, M, tolerance, maximum_iterations, variant, rvlen, maxit )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      n_omega = 0; dimU0 = 0; hr = 0; hc = 0
      pom = c_null_ptr; pu = c_null_ptr; ph = c_null_ptr
      if ( present(omega) ) then
         n_omega = size(omega)
         allocate( bom(max(n_omega,1)) )
         pom = c_loc(bom)
      end if
      if ( present(U0) ) then
         if ( allocated(U0) ) then
            dimU0 = size(U0,2)
            allocate( bU0(n, max(dimU0,1)) )
            pu = c_loc(bU0)
         end if
      end if
      if ( present(H) ) then
         hr = size(H,1); hc = size(H,2)
         allocate( bH(max(hr,1), max(hc,1)) )
         ph = c_loc(bH)
      end if
      bb = b
      call idrs_check( idrs_idrs_ritz_c128(matrix_handle(A, .true.), n, bb, s, o, x, st, pom, n_omega, pu, dimU0, &
                                           ph, hr, hc, nritz, converged), 'IDRS_RITZ' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( n_omega > 0 ) omega = bom(1:n_omega)
      if ( dimU0 > 0 ) U0 = bU0
      if ( present(H) ) H = bH(1:hr,1:hc)
   end function CIDRS_RITZ

! RQR_eig (dense_la_module.f90:295-403): eigenvalues of H(1:n,1:n),
! n = size(H,2); U: Schur vectors
   function RQR_eig( H, U, T, converged ) result(eig)
      real(kind=rp), intent(in)               :: H(:,:)
      complex(kind=cp), intent(out), optional, target :: U(size(H,2),size(H,2))
      complex(kind=cp), intent(out), optional :: T(size(H,2),size(H,2))
      integer, optional                       :: converged
      complex(kind=cp)                        :: eig(size(H,2))
      complex(kind=cp), allocatable           :: Hc(:,:)
      integer(c_int)                          :: conv
      type(c_ptr)                             :: pu

      if ( present(T) ) error stop 'QR_eig: the Schur form T is not offered'
      Hc = cmplx(H, 0._rp, kind=cp)
      pu = c_null_ptr
      if ( present(U) ) pu = c_loc(U)
      call idrs_check( idrs_qr_eig(size(H,2), Hc, size(H,1), 0, eig, pu, conv), 'QR_eig' )
      if ( present(converged) ) converged = conv
   end function RQR_eig

! CQR_eig (dense_la_module.f90:192-293)
   function CQR_eig( H, U, T, converged ) result(eig)
      complex(kind=cp), intent(in)            :: H(:,:)
      complex(kind=cp), intent(out), optional, target :: U(size(H,2),size(H,2))
      complex(kind=cp), intent(out), optional :: T(size(H,2),size(H,2))
      integer, optional                       :: converged
      complex(kind=cp)                        :: eig(size(H,2))
      complex(kind=cp), allocatable           :: Hc(:,:)
      integer(c_int)                          :: conv
      type(c_ptr)                             :: pu

      if ( present(T) ) error stop 'QR_eig: the Schur form T is not offered'
      Hc = H
      pu = c_null_ptr
      if ( present(U) ) pu = c_loc(U)
      call idrs_check( idrs_qr_eig(size(H,2), Hc, size(H,1), 1, eig, pu, conv), 'QR_eig' )
      if ( present(converged) ) converged = conv
   end function CQR_eig

! RSI_power (dense_la_module.f90:532-603)
   function RSI_power( H, s, shift ) result(schurvec)
      real(kind=rp), intent(in)               :: H(:,:)
      integer, intent(in)                     :: s
      real(kind=rp), optional, intent(in)     :: shift
      real(kind=rp)                           :: schurvec(size(H,2),s)
      complex(kind=cp), allocatable           :: Hc(:,:)
      complex(kind=cp)                        :: sh, v(size(H,2),s)

      Hc = cmplx(H, 0._rp, kind=cp)
      sh = (0._cp, 0._cp)
      if ( present(shift) ) sh = cmplx(shift, 0._rp, kind=cp)
      call idrs_check( idrs_si_power(size(H,2), Hc, size(H,1), 0, s, sh, 1_c_int64_t, 0_c_int64_t, v), 'SI_power' )
      schurvec = real(v, kind=rp)
   end function RSI_power

! CSI_power (dense_la_module.f90:455-530)
   function CSI_power( H, s, shift ) result(schurvec)
      complex(kind=cp), intent(in)            :: H(:,:)
      integer, intent(in)                     :: s
      complex(kind=cp), optional, intent(in)  :: shift
      complex(kind=cp)                        :: schurvec(size(H,2),s)
      complex(kind=cp), allocatable           :: Hc(:,:)
      complex(kind=cp)                        :: sh

      Hc = H
      sh = (0._cp, 0._cp)
      if ( present(shift) ) sh = shift
      call idrs_check( idrs_si_power(size(H,2), Hc, size(H,1), 1, s, sh, 1_c_int64_t, 0_c_int64_t, schurvec), &
                       'SI_power' )
   end function CSI_power

end module ritz_module
