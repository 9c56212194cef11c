!
! idrs_module: drop-in for the reference's MODULES/IDRS/idrs_module.f90
! solver interface.  x = IDRS(A, b, s, ...) keeps the reference's generic
! name, argument order, optional arguments and flags; the iteration runs on
! the GPU in libidrs_hip.so (device-resident state, fused kernels).
!
!   IDRS     (RIDRS, CIDRS)          idrs_module.f90:107-251, idrs.inc:1-311
!   QMRIDR   (RQMRIDR, CQMRIDR)      idrs_module.f90:253-408, qmridr.inc:1-263
!   MSQMRIDR (RMSQMRIDR, CMSQMRIDR)  idrs_module.f90:410-570, msqmridr.inc:1-293
!   MSIDRS   (RMSIDRS, CMSIDRS)      idrs_module.f90:572-713, msidrs.inc:1-334
!   NORM, INPROD               idrs_module.f90:787-837 (host arrays)
!
! The optional output H of IDRS (Hessenberg matrix, idrs.inc:81-90) is
! filled on the device; IDRS_RITZ is in ritz_module.
!
module idrs_module

   use precision_module
   use matrix_module
   use idrs_capi
   use iso_c_binding

   implicit none

   private
   public :: IDRS, QMRIDR, MSIDRS, MSQMRIDR, NORM, INPROD

   interface IDRS
      module procedure CIDRS, RIDRS
   end interface

   interface QMRIDR
      module procedure CQMRIDR, RQMRIDR
   end interface

   interface MSIDRS
      module procedure CMSIDRS, RMSIDRS
   end interface

   interface MSQMRIDR
      module procedure CMSQMRIDR, RMSQMRIDR
   end interface

   interface NORM
      module procedure CNORM, RNORM
   end interface

   interface INPROD
      module procedure CINPROD, RINPROD
   end interface

contains

! Fills the options common to RIDRS and CIDRS (presence flags as present()).
   subroutine common_opts( o, n, s, M1, cplx, tolerance, maximum_iterations, variant, resvec_len, maxit )
      type(idrs_opts_t), intent(inout)       :: o
      integer, intent(in)                    :: n, s
      type(matrix), optional, intent(in)     :: M1
      logical, intent(in)                    :: cplx
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, intent(in)                    :: resvec_len
      integer, intent(out)                   :: maxit

! v/M1 divides by M1%P's diagonal; a default type(matrix) is the identity
      if ( present(M1) ) then
         o%has_M1 = 1
         if ( M1%P%which_precon == 1 ) o%M1 = M1%P%hp
      end if
      if ( present(tolerance) ) then
         if ( tolerance < 0 ) error stop "Illegal value for parameter tolerance"
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
      if ( s < 1 ) error stop "IDRS: s must be positive"
   end subroutine common_opts

! Fills the options of QMRIDR, MSIDRS and MSQMRIDR.  nsig is the number of
! shifts (1 for QMRIDR); resvec_rows/resvec_cols are the shape of resvec
! (-1 when absent).
   subroutine krylov_opts( o, n, s, nsig, M1, tolerance, maximum_iterations, variant, &
                           inner_s, inner_tolerance, inner_iterations, resvec_rows, resvec_cols )
      type(idrs_krylov_opts_t), intent(inout) :: o
      integer, intent(in)                    :: n, s, nsig
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations, variant
      integer, optional, intent(in)          :: inner_s
      real(kind=rp), optional, intent(in)    :: inner_tolerance
      integer, optional, intent(in)          :: inner_iterations
      integer, intent(in)                    :: resvec_rows, resvec_cols
      integer                                :: maxit

      if ( present(M1) ) then
         o%has_M1 = 1
         if ( M1%P%which_precon == 1 ) o%M1 = M1%P%hp
      end if
      if ( present(tolerance) ) then
         if ( tolerance < 0 ) error stop "Illegal value for parameter tolerance"
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
      if ( present(inner_s) ) then
         o%has_inner_s = 1
         o%inner_s = inner_s
      end if
      if ( present(inner_tolerance) ) then
         o%has_inner_tolerance = 1
         o%inner_tolerance = inner_tolerance
      end if
      if ( present(inner_iterations) ) then
         o%has_inner_iterations = 1
         o%inner_iterations = inner_iterations
      end if
      if ( resvec_rows >= 0 ) then
         if ( maxit+1 > resvec_rows ) &
            error stop "Length of vector with residual norms too small, should be maxit+1"
         if ( resvec_cols /= nsig ) error stop "resvec must have one column per shift"
      end if
      if ( s < 1 ) error stop "s must be positive"
   end subroutine krylov_opts

   function RIDRS( A, b, s, &                                    ! required
                   M1, tolerance, maximum_iterations, variant, & ! optional input
                   flag, relres, iterations, &                   ! optional output
                   x0, U0, omega, resvec, H ) result(x)          ! optional arrays
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)              :: b(:)
      integer, intent(in)                    :: s
      real(kind=rp)                          :: x(size(b))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres
      integer, optional, intent(out)         :: iterations
      real(kind=rp), optional, intent(in)    :: x0(:)
      real(kind=rp), optional, allocatable, &
                               intent(in)    :: U0(:,:)
      real(kind=rp), optional, intent(in)    :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:)
      real(kind=cp), optional, intent(out)   :: H(:,:)

      type(idrs_opts_t)                      :: o
      type(idrs_stats_t)                     :: st
      real(kind=rp), allocatable, target     :: bx0(:), bU0(:,:), bom(:), brv(:), bH(:,:)
      real(kind=rp), allocatable             :: bb(:)
      integer                                :: n, maxit, rvlen, nritz

      n = size(b)
      nritz = 0
      if ( present(H) ) then
         nritz = size(H,1) - 1
         if ( nritz > 0 ) then
            if ( size(H,2) /= nritz ) stop "Second dimension of H incompatible, with first"
            allocate( bH(nritz+1, nritz) )
            o%H = c_loc(bH)
            o%H_nritz = nritz
         end if
      end if
      rvlen = -1
      if ( present(resvec) ) rvlen = size(resvec)
      call common_opts( o, n, s, M1, .false., tolerance, maximum_iterations, variant, rvlen, maxit )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(U0) ) then
         if ( allocated(U0) ) then
            if ( size(U0) == n*s ) then
               bU0 = U0
               o%U0 = c_loc(bU0)
            end if
         end if
      end if
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      bb = b
      call idrs_check( idrs_idrs_f64(matrix_handle(A, .false.), n, bb, s, o, x, st), 'IDRS' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( nritz > 0 )           H = bH
   end function RIDRS

   function CIDRS( A, b, s, &                                     ! required
                   M1, tolerance, maximum_iterations, variant, &  ! optional input
                   flag, relres, iterations, &                    ! optional output
                   x0, U0, omega, resvec, H ) result(x)           ! optional arrays
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)           :: b(:)
      integer, intent(in)                    :: s
      complex(kind=cp)                       :: x(size(b))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres
      integer, optional, intent(out)         :: iterations
      complex(kind=cp), optional, intent(in) :: x0(:)
      complex(kind=cp), optional, allocatable, &
                               intent(in)    :: U0(:,:)
      complex(kind=cp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:)
      complex(kind=cp), optional, intent(out):: H(:,:)

      type(idrs_opts_t)                      :: o
      type(idrs_stats_t)                     :: st
      complex(kind=cp), allocatable, target  :: bx0(:), bU0(:,:), bom(:), bH(:,:)
      real(kind=rp), allocatable, target     :: brv(:)
      complex(kind=cp), allocatable          :: bb(:)
      integer                                :: n, maxit, rvlen, nritz

      n = size(b)
      nritz = 0
      if ( present(H) ) then
         nritz = size(H,1) - 1
         if ( nritz > 0 ) then
            if ( size(H,2) /= nritz ) stop "Second dimension of H incompatible, with first"
            allocate( bH(nritz+1, nritz) )
            o%H = c_loc(bH)
            o%H_nritz = nritz
         end if
      end if
      rvlen = -1
      if ( present(resvec) ) rvlen = size(resvec)
      call common_opts( o, n, s, M1, .true., tolerance, maximum_iterations, variant, rvlen, maxit )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(U0) ) then
         if ( allocated(U0) ) then
            if ( size(U0) == n*s ) then
               bU0 = U0
               o%U0 = c_loc(bU0)
            end if
         end if
      end if
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      bb = b
      call idrs_check( idrs_idrs_c128(matrix_handle(A, .true.), n, bb, s, o, x, st), 'IDRS' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( nritz > 0 )           H = bH
   end function CIDRS

   function RQMRIDR( A, b, s, &                                    ! required
                    M1, tolerance, maximum_iterations, &          ! optional input
                    flag, relres, iterations, &                   ! optional output
                    inner_s, inner_tolerance, inner_iterations, & ! optional input
                    x0, omega, resvec ) result(x)                 ! optional arrays
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)          :: b(:)
      integer, intent(in)                    :: s
      real(kind=rp)                          :: x(size(b))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres
      integer, optional, intent(out)         :: iterations
      integer, optional, intent(in)          :: inner_s
      real(kind=rp), optional, intent(in)    :: inner_tolerance
      integer, optional, intent(in)          :: inner_iterations
      real(kind=rp), optional, intent(in) :: x0(:)
      real(kind=rp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      real(kind=rp), allocatable, target :: bx0(:), bom(:)
      real(kind=rp), allocatable, target     :: brv(:)
      real(kind=rp), allocatable :: bb(:)
      integer                                :: n, rvrows

      n = size(b)
      rvrows = -1
      if ( present(resvec) ) rvrows = size(resvec)
      call krylov_opts( o, n, s, 1, M1, tolerance, maximum_iterations, &
                        inner_s=inner_s, inner_tolerance=inner_tolerance, inner_iterations=inner_iterations, &
                        resvec_rows=rvrows, resvec_cols=1 )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      bb = b
      call idrs_check( idrs_qmridr_f64(matrix_handle(A, .false.), n, bb, s, o, x, st), 'QMRIDR' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
   end function RQMRIDR

   function CQMRIDR( A, b, s, &                                    ! required
                    M1, tolerance, maximum_iterations, &          ! optional input
                    flag, relres, iterations, &                   ! optional output
                    inner_s, inner_tolerance, inner_iterations, & ! optional input
                    x0, omega, resvec ) result(x)                 ! optional arrays
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)       :: b(:)
      integer, intent(in)                    :: s
      complex(kind=cp)                       :: x(size(b))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres
      integer, optional, intent(out)         :: iterations
      integer, optional, intent(in)          :: inner_s
      real(kind=rp), optional, intent(in)    :: inner_tolerance
      integer, optional, intent(in)          :: inner_iterations
      complex(kind=cp), optional, intent(in) :: x0(:)
      complex(kind=cp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      complex(kind=cp), allocatable, target :: bx0(:), bom(:)
      real(kind=rp), allocatable, target     :: brv(:)
      complex(kind=cp), allocatable :: bb(:)
      integer                                :: n, rvrows

      n = size(b)
      rvrows = -1
      if ( present(resvec) ) rvrows = size(resvec)
      call krylov_opts( o, n, s, 1, M1, tolerance, maximum_iterations, &
                        inner_s=inner_s, inner_tolerance=inner_tolerance, inner_iterations=inner_iterations, &
                        resvec_rows=rvrows, resvec_cols=1 )
      if ( present(x0) ) then
         bx0 = x0
         o%x0 = c_loc(bx0)
      end if
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(size(resvec)) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = size(resvec)
      end if
      bb = b
      call idrs_check( idrs_qmridr_c128(matrix_handle(A, .true.), n, bb, s, o, x, st), 'QMRIDR' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = st%relres
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
   end function CQMRIDR

   function RMSIDRS( A, b, sigma, s, &                             ! required
                    M1, tolerance, maximum_iterations, variant, & ! optional input
                    flag, relres, iterations, &                   ! optional output
                    omega, resvec, colfac ) result(x)             ! optional arrays
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)  :: b(:)
      real(kind=rp), intent(in)  :: sigma(:)
      integer, intent(in)                    :: s
      real(kind=rp)           :: x(size(b),size(sigma))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres(size(sigma))
      integer, optional, intent(out)         :: iterations
      real(kind=rp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:,:)
      real(kind=rp), optional, intent(out) :: colfac(:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      real(kind=rp), allocatable, target :: bom(:)
      real(kind=rp), allocatable, target :: bcf(:)
      real(kind=rp), allocatable, target     :: brv(:,:)
      real(kind=rp), allocatable :: bb(:), bsig(:)
      real(kind=rp)                          :: rr(size(sigma))
      integer                                :: n, nsig, rvrows, rvcols

      n = size(b)
      nsig = size(sigma)
      rvrows = -1
      rvcols = 0
      if ( present(resvec) ) then
         rvrows = size(resvec,1)
         rvcols = size(resvec,2)
      end if
      call krylov_opts( o, n, s, nsig, M1, tolerance, maximum_iterations, variant=variant, &
                        resvec_rows=rvrows, resvec_cols=rvcols )
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(rvrows, rvcols) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = rvrows
      end if
      if ( present(colfac) ) then
         if ( size(colfac) < nsig ) error stop "colfac must hold one value per shift"
         allocate( bcf(nsig) )
         o%colfac = c_loc(bcf)
      end if
      bb = b
      bsig = sigma
      call idrs_check( idrs_msidrs_f64(matrix_handle(A, .false.), n, bb, nsig, bsig, s, o, x, rr, st), &
                       'MSIDRS' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = rr
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( present(colfac) )     colfac(1:nsig) = bcf
   end function RMSIDRS

   function CMSIDRS( A, b, sigma, s, &                             ! required
                    M1, tolerance, maximum_iterations, variant, & ! optional input
                    flag, relres, iterations, &                   ! optional output
                    omega, resvec, colfac ) result(x)             ! optional arrays
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)  :: b(:)
      complex(kind=cp), intent(in)  :: sigma(:)
      integer, intent(in)                    :: s
      complex(kind=cp)           :: x(size(b),size(sigma))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: variant
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres(size(sigma))
      integer, optional, intent(out)         :: iterations
      complex(kind=cp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:,:)
      complex(kind=cp), optional, intent(out) :: colfac(:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      complex(kind=cp), allocatable, target :: bom(:)
      complex(kind=cp), allocatable, target :: bcf(:)
      real(kind=rp), allocatable, target     :: brv(:,:)
      complex(kind=cp), allocatable :: bb(:), bsig(:)
      real(kind=rp)                          :: rr(size(sigma))
      integer                                :: n, nsig, rvrows, rvcols

      n = size(b)
      nsig = size(sigma)
      rvrows = -1
      rvcols = 0
      if ( present(resvec) ) then
         rvrows = size(resvec,1)
         rvcols = size(resvec,2)
      end if
      call krylov_opts( o, n, s, nsig, M1, tolerance, maximum_iterations, variant=variant, &
                        resvec_rows=rvrows, resvec_cols=rvcols )
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(rvrows, rvcols) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = rvrows
      end if
      if ( present(colfac) ) then
         if ( size(colfac) < nsig ) error stop "colfac must hold one value per shift"
         allocate( bcf(nsig) )
         o%colfac = c_loc(bcf)
      end if
      bb = b
      bsig = sigma
      call idrs_check( idrs_msidrs_c128(matrix_handle(A, .true.), n, bb, nsig, bsig, s, o, x, rr, st), &
                       'MSIDRS' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = rr
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
      if ( present(colfac) )     colfac(1:nsig) = bcf
   end function CMSIDRS

   function RMSQMRIDR( A, b, sigma, s, &                             ! required
                    M1, tolerance, maximum_iterations, &          ! optional input
                    flag, relres, iterations, &                   ! optional output
                    inner_s, inner_tolerance, inner_iterations, & ! optional input
                    omega, resvec ) result(x)                     ! optional arrays
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)  :: b(:)
      real(kind=rp), intent(in)  :: sigma(:)
      integer, intent(in)                    :: s
      real(kind=rp)           :: x(size(b),size(sigma))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: inner_s
      real(kind=rp), optional, intent(in)    :: inner_tolerance
      integer, optional, intent(in)          :: inner_iterations
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres(size(sigma))
      integer, optional, intent(out)         :: iterations
      real(kind=rp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:,:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      real(kind=rp), allocatable, target :: bom(:)
      real(kind=rp), allocatable, target     :: brv(:,:)
      real(kind=rp), allocatable :: bb(:), bsig(:)
      real(kind=rp)                          :: rr(size(sigma))
      integer                                :: n, nsig, rvrows, rvcols

      n = size(b)
      nsig = size(sigma)
      rvrows = -1
      rvcols = 0
      if ( present(resvec) ) then
         rvrows = size(resvec,1)
         rvcols = size(resvec,2)
      end if
      call krylov_opts( o, n, s, nsig, M1, tolerance, maximum_iterations, inner_s=inner_s, inner_tolerance=inner_tolerance, inner_iterations=inner_iterations, &
                        resvec_rows=rvrows, resvec_cols=rvcols )
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(rvrows, rvcols) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = rvrows
      end if
      bb = b
      bsig = sigma
      call idrs_check( idrs_msqmridr_f64(matrix_handle(A, .false.), n, bb, nsig, bsig, s, o, x, rr, st), &
                       'MSQMRIDR' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = rr
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
   end function RMSQMRIDR

   function CMSQMRIDR( A, b, sigma, s, &                             ! required
                    M1, tolerance, maximum_iterations, &          ! optional input
                    flag, relres, iterations, &                   ! optional output
                    inner_s, inner_tolerance, inner_iterations, & ! optional input
                    omega, resvec ) result(x)                     ! optional arrays
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)  :: b(:)
      complex(kind=cp), intent(in)  :: sigma(:)
      integer, intent(in)                    :: s
      complex(kind=cp)           :: x(size(b),size(sigma))
      type(matrix), optional, intent(in)     :: M1
      real(kind=rp), optional, intent(in)    :: tolerance
      integer, optional, intent(in)          :: maximum_iterations
      integer, optional, intent(in)          :: inner_s
      real(kind=rp), optional, intent(in)    :: inner_tolerance
      integer, optional, intent(in)          :: inner_iterations
      integer, optional, intent(out)         :: flag
      real(kind=rp), optional, intent(out)   :: relres(size(sigma))
      integer, optional, intent(out)         :: iterations
      complex(kind=cp), optional, intent(in) :: omega(:)
      real(kind=rp), optional, intent(out)   :: resvec(:,:)

      type(idrs_krylov_opts_t)               :: o
      type(idrs_stats_t)                     :: st
      complex(kind=cp), allocatable, target :: bom(:)
      real(kind=rp), allocatable, target     :: brv(:,:)
      complex(kind=cp), allocatable :: bb(:), bsig(:)
      real(kind=rp)                          :: rr(size(sigma))
      integer                                :: n, nsig, rvrows, rvcols

      n = size(b)
      nsig = size(sigma)
      rvrows = -1
      rvcols = 0
      if ( present(resvec) ) then
         rvrows = size(resvec,1)
         rvcols = size(resvec,2)
      end if
      call krylov_opts( o, n, s, nsig, M1, tolerance, maximum_iterations, inner_s=inner_s, inner_tolerance=inner_tolerance, inner_iterations=inner_iterations, &
                        resvec_rows=rvrows, resvec_cols=rvcols )
      if ( present(omega) ) then
         if ( size(omega) >= 1 ) then
            bom = omega
            o%omega = c_loc(bom)
            o%n_omega = size(omega)
         end if
      end if
      if ( present(resvec) ) then
         allocate( brv(rvrows, rvcols) )
         brv = 0._rp
         o%resvec = c_loc(brv)
         o%resvec_len = rvrows
      end if
      bb = b
      bsig = sigma
      call idrs_check( idrs_msqmridr_c128(matrix_handle(A, .true.), n, bb, nsig, bsig, s, o, x, rr, st), &
                       'MSQMRIDR' )
      if ( present(flag) )       flag = st%flag
      if ( present(relres) )     relres = rr
      if ( present(iterations) ) iterations = st%iterations
      if ( present(resvec) )     resvec = brv
   end function CMSQMRIDR

! Norm and inner product of vectors held by the caller (one process: the
! reference's co_sum is the identity).
   function RNORM( w )
      real(kind=rp), intent(in)                 :: w(:)
      real(kind=rp)                             :: RNORM
      RNORM = sqrt( sum( w**2 ) )
   end function RNORM

   function CNORM( w )
      complex(kind=cp), intent(in)              :: w(:)
      real(kind=rp)                             :: CNORM
      CNORM = sqrt( sum( abs(w)**2 ) )
   end function CNORM

   function RINPROD( u, v )
      real(kind=rp), intent(in)                 :: u(:), v(:)
      real(kind=rp)                             :: RINPROD
      RINPROD = dot_product( u, v )
   end function RINPROD

   function CINPROD( u, v )
      complex(kind=cp), intent(in)              :: u(:), v(:)
      complex(kind=cp)                          :: CINPROD
      CINPROD = dot_product( u, v )
   end function CINPROD

end module idrs_module
