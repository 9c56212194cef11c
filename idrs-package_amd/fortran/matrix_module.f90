!
! matrix_module: drop-in for the reference's MODULES/PP_MATRIX/matrix_module.f90
! public interface on the hot path.  type(matrix) keeps the reference's
! components (K, M, C, P, seed, polynomial coefficients, flags), so code
! that fills them as PP_IDRS/real_matrix.inc does keeps working.  Every
! product runs in libidrs_hip.so: the Fortran state is turned into a device
! operator (diagonal scaling folded into the matrix values, Horner
! polynomial of SpMV+axpy kernels) that is cached per state.
!
!   operator(*)   rmatvec / cmatvec          matrix_module.f90:135-167
!   operator(/)   rprecon / cprecon          matrix_module.f90:306-400
!   real_preconditioner / complex_preconditioner  :551-722
!   CHEBYSHEV / NEUMANN / COMPUTE_ALPHA      matrix_module.f90:851-1057
!   SCALEBACK     rscaleback / cscaleback    matrix_module.f90:808-849
!                 ms_rscaleback / ms_cscaleback               :724-806
!   MS_PARAMETERS real_ / complex_ms_parameters               :402-549
!   Pmul          deflation projectors        matrix_module.f90:1369-1431
!   make_deflation / add_deflation            matrix_module.f90:1284-1623
!   real_mv / complex_mv (K - sM, K + sC + s^2 M, block systems) :169-264
!   filter_ritzval, ANALYSE_SPECTRUM         matrix_module.f90:1059-1282
!                 (host routines of the library)
!
! Mass (A%M) and damping (A%C) matrices, block systems and deflation run on
! the device operator; make_deflation also fills A%rAV/A%rVtAV (A%cAV,
! A%cVtAV) as the reference does.  Shift-and-invert (A%shift_invert with
! A%rshift / A%cshift and an LU preconditioner A%P = REAL_PRECON(2|3, ...))
! runs on the device operator too: the block system's inverse, the
! multishift parameters and scale-back factor.
!
module matrix_module

   use precision_module
   use user_module
   use idrs_capi
   use iso_c_binding

   implicit none

   private
   public :: matrix, operator(*), operator(/), CHEBYSHEV, NEUMANN, SCALEBACK, COMPUTE_ALPHA, Pmul, &
             real_preconditioner, complex_preconditioner, matrix_handle, MS_PARAMETERS, &
             filter_ritzval, ANALYSE_SPECTRUM, make_deflation, add_deflation, real_mv, complex_mv

   type :: matrix
      type(user_matrix)              :: K, M, C, P
      real(kind=rp)                  :: rseed = 0._rp
      complex(kind=cp)               :: cseed = (0._cp,0._cp)
      real(kind=rp)                  :: rshift = 0._rp
      complex(kind=cp)               :: cshift = (0._cp,0._cp)
      integer                        :: preconditioner = 0
      integer                        :: degree = 0
      real(kind=rp), allocatable     :: ralpha_p(:)
      real(kind=rp), allocatable     :: ralpha_s(:,:)
      complex(kind=cp), allocatable  :: calpha_p(:)
      complex(kind=cp), allocatable  :: calpha_s(:,:)
      integer                        :: n_deflation = 0
      real(kind=rp), allocatable     :: rV(:,:), rAV(:,:), rVtAV(:,:)
      complex(kind=cp), allocatable  :: cV(:,:), cAV(:,:), cVtAV(:,:)
! which make_deflation call V comes from (device operators are cached per
! deflation space)
      integer                        :: deflation_id = 0
      logical                        :: mass_matrix = .false.
      logical                        :: damping_matrix = .false.
      logical                        :: multishift = .false.
      logical                        :: block_system = .false.
      logical                        :: shift_invert = .false.
   end type matrix

   INTERFACE OPERATOR(*)
      module procedure rmatvec, cmatvec
   END INTERFACE

   INTERFACE OPERATOR(/)
      module procedure rprecon, cprecon
   END INTERFACE

   interface CHEBYSHEV
      module procedure RCHEBYSHEV, CCHEBYSHEV
   end interface

   interface NEUMANN
      module procedure RNEUMANN, CNEUMANN
   end interface

   interface SCALEBACK
      module procedure CSCALEBACK, RSCALEBACK, MS_CSCALEBACK, MS_RSCALEBACK
   end interface

   interface COMPUTE_ALPHA
      module procedure CCOMPUTE_ALPHA, RCOMPUTE_ALPHA
   end interface

   interface Pmul
      module procedure rPmul, cPmul
   end interface

   interface MS_PARAMETERS
      module procedure COMPLEX_MS_PARAMETERS, REAL_MS_PARAMETERS
   end interface

   interface ANALYSE_SPECTRUM
      module procedure RANALYSE_SPECTRUM, CANALYSE_SPECTRUM
   end interface

   interface make_deflation
      module procedure rmake_deflation, cmake_deflation
   end interface

   interface add_deflation
      module procedure radd_deflation, cadd_deflation, rms_add_deflation, cms_add_deflation
   end interface

! Cache of device operators built from type(matrix) states
   integer, parameter :: NCACHE = 8
   type :: op_entry
      type(c_ptr)                   :: op = c_null_ptr
      type(c_ptr)                   :: k = c_null_ptr
      type(c_ptr)                   :: m = c_null_ptr, c = c_null_ptr
      type(c_ptr)                   :: pk = c_null_ptr, pm = c_null_ptr, pc = c_null_ptr
      logical                       :: block = .false., multishift = .false., si = .false.
      complex(kind=cp)              :: si_shift = (0._cp,0._cp)
      integer                       :: defl_id = 0
      integer                       :: which_precon = 0
      logical                       :: has_shift = .false.
      complex(kind=cp)              :: shift = (0._cp,0._cp)
      complex(kind=cp)              :: seed = (0._cp,0._cp)
      logical                       :: cplx = .false.
      integer                       :: degree = 0
      complex(kind=cp), allocatable :: alpha(:)
      integer                       :: stamp = 0
   end type op_entry
   type(op_entry), save :: cache(NCACHE)
   integer, save        :: clock = 0
   integer, save        :: deflation_count = 0

contains

! Device operator for the current state of A (cplx: complex arithmetic).
   function matrix_handle( A, cplx ) result(h)
      type(matrix), intent(in)          :: A
      logical, intent(in)               :: cplx
      type(c_ptr)                       :: h
      complex(kind=cp)                  :: seed, shift
      complex(kind=cp), allocatable     :: alpha(:)
      integer                           :: i, victim, deg, did
      integer(c_int)                    :: status
      type(c_ptr)                       :: mh, ch, pk, pm, pc
      logical                           :: own_diag
      real(c_double)                    :: sh(2)
      complex(kind=cp)                  :: sis

      if ( .not. c_associated(A%K%h) ) error stop 'matrix: the system matrix A%K is not set'
      mh = c_null_ptr
      ch = c_null_ptr
      if ( A%mass_matrix ) mh = A%M%h
      if ( A%damping_matrix ) ch = A%C%h
      did = 0
      if ( A%n_deflation > 0 ) did = A%deflation_id
      if ( A%n_deflation > 0 .and. did == 0 ) error stop 'matrix: n_deflation is set, call make_deflation'
      pk = c_null_ptr
      pm = c_null_ptr
      pc = c_null_ptr
      if ( A%P%which_precon == 1 ) then
         pk = A%P%h
         pm = A%P%hm
         pc = A%P%hc
      elseif ( A%P%which_precon == 2 .or. A%P%which_precon == 3 ) then
! the LU factors live in the preconditioner's own operator: shared
         pk = A%P%hp
      end if
      sis = A%cshift
      if ( .not. cplx ) sis = cmplx(A%rshift, 0._rp, kind=cp)
      if ( cplx .and. A%rshift /= 0._rp .and. A%cshift == (0._cp,0._cp) ) sis = cmplx(A%rshift, 0._rp, kind=cp)
! a diagonal preconditioner built from A%K alone folds into the matrix
! values; any other (e.g. a lumped mass matrix) divides by its diagonal
      own_diag = A%P%which_precon == 1 .and. c_associated(pk, A%K%h) .and. .not. c_associated(pm) &
                 .and. .not. c_associated(pc)

      if ( cplx ) then
         seed = A%cseed
         if ( A%rseed /= 0._rp .and. A%cseed == (0._cp,0._cp) ) seed = cmplx(A%rseed, 0._rp, kind=cp)
         shift = A%P%cshift
         if ( .not. A%P%is_complex ) shift = cmplx(A%P%rshift, 0._rp, kind=cp)
      else
         seed = cmplx(A%rseed, 0._rp, kind=cp)
         shift = cmplx(A%P%rshift, 0._rp, kind=cp)
      end if
      deg = A%degree
      allocate( alpha(0:deg) )
      alpha = (0._cp,0._cp)
      if ( deg > 0 ) then
         if ( cplx .and. allocated(A%calpha_p) ) then
            alpha = A%calpha_p(0:deg)
         elseif ( allocated(A%ralpha_p) ) then
            alpha = cmplx(A%ralpha_p(0:deg), 0._rp, kind=cp)
         elseif ( allocated(A%calpha_p) ) then
            alpha = A%calpha_p(0:deg)
         else
            deg = 0
         end if
      end if

      clock = clock + 1
      do i = 1, NCACHE
         if ( .not. c_associated(cache(i)%op) ) cycle
         if ( .not. c_associated(cache(i)%k, A%K%h) ) cycle
         if ( .not. same(cache(i)%m, mh) .or. .not. same(cache(i)%c, ch) ) cycle
         if ( .not. same(cache(i)%pk, pk) .or. .not. same(cache(i)%pm, pm) .or. .not. same(cache(i)%pc, pc) ) cycle
         if ( (cache(i)%block .neqv. A%block_system) .or. (cache(i)%multishift .neqv. A%multishift) ) cycle
         if ( cache(i)%defl_id /= did ) cycle
         if ( (cache(i)%si .neqv. A%shift_invert) .or. cache(i)%si_shift /= sis ) cycle
         if ( cache(i)%which_precon /= A%P%which_precon .or. (cache(i)%has_shift .neqv. A%P%has_shift) ) cycle
         if ( cache(i)%shift /= shift .or. cache(i)%seed /= seed .or. (cache(i)%cplx .neqv. cplx) ) cycle
         if ( cache(i)%degree /= deg ) cycle
         if ( any(cache(i)%alpha(0:deg) /= alpha(0:deg)) ) cycle
         cache(i)%stamp = clock
         h = cache(i)%op
         return
      end do

! build a new operator, replacing the least recently used entry
      victim = 1
      do i = 2, NCACHE
         if ( cache(i)%stamp < cache(victim)%stamp ) victim = i
      end do
      if ( c_associated(cache(victim)%op) ) status = idrs_matrix_destroy(cache(victim)%op)
      cache(victim)%op = c_null_ptr
      call idrs_check( idrs_matrix_create(A%K%h, h), 'matrix' )
      if ( cplx ) then
         call idrs_check( idrs_matrix_set_seed_c128(h, seed), 'matrix' )
      else
         call idrs_check( idrs_matrix_set_seed_f64(h, real(seed, kind=rp)), 'matrix' )
      end if
      if ( c_associated(mh) ) call idrs_check( idrs_matrix_set_mass(h, mh), 'matrix' )
      if ( c_associated(ch) ) call idrs_check( idrs_matrix_set_damping(h, ch), 'matrix' )
      call idrs_check( idrs_matrix_set_flags(h, merge(1, 0, A%block_system), merge(1, 0, A%multishift)), 'matrix' )
      if ( own_diag ) then
         if ( A%P%has_shift ) then
            if ( cplx ) then
               call idrs_check( idrs_matrix_set_precon_c128(h, 1, shift), 'matrix' )
            else
               call idrs_check( idrs_matrix_set_precon_f64(h, 1, real(shift, kind=rp)), 'matrix' )
            end if
         else
            call idrs_check( idrs_matrix_set_precon_f64(h, 1), 'matrix' )
         end if
      elseif ( A%P%which_precon == 2 .or. A%P%which_precon == 3 ) then
         call idrs_check( idrs_matrix_share_lu(h, A%P%hp), 'matrix' )
      elseif ( A%P%which_precon == 1 ) then
         sh = [ real(shift, kind=rp), aimag(shift) ]
         if ( A%P%has_shift ) then
            call idrs_check( idrs_matrix_set_precon_diag(h, pk, pm, pc, sh, merge(1, 0, cplx)), 'matrix' )
         else
            call idrs_check( idrs_matrix_set_precon_diag(h, pk, pm, pc, is_complex=merge(1, 0, cplx)), 'matrix' )
         end if
      end if
      if ( A%shift_invert ) then
         sh = [ real(sis, kind=rp), aimag(sis) ]
         call idrs_check( idrs_matrix_set_shift_invert(h, 1, sh, merge(1, 0, cplx)), 'matrix' )
      end if
      call idrs_check( idrs_matrix_set_alpha(h, deg, alpha), 'matrix' )
      if ( did > 0 ) then
         if ( cplx ) then
            if ( allocated(A%cV) ) then
               call idrs_check( idrs_make_deflation_c128(h, size(A%cV,2), A%cV), 'make_deflation' )
            else
               call idrs_check( idrs_make_deflation_c128(h, size(A%rV,2), cmplx(A%rV, 0._rp, kind=cp)), &
                                'make_deflation' )
            end if
         else
            if ( .not. allocated(A%rV) ) error stop 'matrix: a complex deflation space on a real operator'
            call idrs_check( idrs_make_deflation_f64(h, size(A%rV,2), A%rV), 'make_deflation' )
         end if
      end if
      cache(victim)%op = h
      cache(victim)%k = A%K%h
      cache(victim)%m = mh
      cache(victim)%c = ch
      cache(victim)%pk = pk
      cache(victim)%pm = pm
      cache(victim)%pc = pc
      cache(victim)%block = A%block_system
      cache(victim)%multishift = A%multishift
      cache(victim)%defl_id = did
      cache(victim)%si = A%shift_invert
      cache(victim)%si_shift = sis
      cache(victim)%which_precon = A%P%which_precon
      cache(victim)%has_shift = A%P%has_shift
      cache(victim)%shift = shift
      cache(victim)%seed = seed
      cache(victim)%cplx = cplx
      cache(victim)%degree = deg
      cache(victim)%alpha = alpha
      cache(victim)%stamp = clock
   end function matrix_handle

   logical function same( a, b )
      type(c_ptr), intent(in) :: a, b
      if ( c_associated(a) ) then
         same = c_associated(a, b)
      else
         same = .not. c_associated(b)
      end if
   end function same

   function real_mv( A, v ) result(w)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      call idrs_check( idrs_system_mv_f64(matrix_handle(A, .false.), v, w), 'real_mv' )
   end function real_mv

   function complex_mv( A, v ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      call idrs_check( idrs_system_mv_c128(matrix_handle(A, .true.), v, w), 'complex_mv' )
   end function complex_mv

   function cmatvec( A, v ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      call idrs_check( idrs_matvec_c128(matrix_handle(A, .true.), v, w), 'matvec' )
   end function cmatvec

   function rmatvec( A, v ) result(w)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      call idrs_check( idrs_matvec_f64(matrix_handle(A, .false.), v, w), 'matvec' )
   end function rmatvec

   function rprecon( v, A ) result(w)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      if ( A%block_system ) then
         call idrs_check( idrs_precon_f64(matrix_handle(A, .false.), v, w), 'precon' )
      else
         w = v/A%P
      end if
   end function rprecon

   function cprecon( v, A ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      if ( A%block_system ) then
         call idrs_check( idrs_precon_c128(matrix_handle(A, .true.), v, w), 'precon' )
      else
         w = v/A%P
      end if
   end function cprecon

   function rPmul( A, v ) result(w)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      if ( A%n_deflation > 0 ) then
         call idrs_check( idrs_pmul_f64(matrix_handle(A, .false.), v, w), 'Pmul' )
      else
         w = v
      end if
   end function rPmul

   function cPmul( A, v ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      if ( A%n_deflation > 0 ) then
         call idrs_check( idrs_pmul_c128(matrix_handle(A, .true.), v, w), 'Pmul' )
      else
         w = v
      end if
   end function cPmul

! make_deflation (matrix_module.f90:1284-1352): keeps V, AV = real_mv(A, V)
! and VtAV = V^T AV (V^H AV) in A; the device operator forms its own copies
   subroutine rmake_deflation( A, V )
      type(matrix), intent(inout)       :: A
      real(kind=rp), intent(in)         :: V(:,:)
      integer                           :: j
      A%n_deflation = size(V,2)
      if ( allocated(A%rV) ) deallocate(A%rV, A%rAV, A%rVtAV)
      if ( allocated(A%cV) ) deallocate(A%cV, A%cAV, A%cVtAV)
      allocate( A%rV(size(V,1),size(V,2)), A%rAV(size(V,1),size(V,2)), A%rVtAV(size(V,2),size(V,2)) )
      A%rV = V
      deflation_count = deflation_count + 1
      A%deflation_id = deflation_count
      do j = 1, size(V,2)
         A%rAV(:,j) = real_mv( A, V(:,j) )
      end do
      A%rVtAV = matmul( transpose(V), A%rAV )
   end subroutine rmake_deflation

   subroutine cmake_deflation( A, V )
      type(matrix), intent(inout)       :: A
      complex(kind=cp), intent(in)      :: V(:,:)
      integer                           :: j
      A%n_deflation = size(V,2)
      if ( allocated(A%rV) ) deallocate(A%rV, A%rAV, A%rVtAV)
      if ( allocated(A%cV) ) deallocate(A%cV, A%cAV, A%cVtAV)
      allocate( A%cV(size(V,1),size(V,2)), A%cAV(size(V,1),size(V,2)), A%cVtAV(size(V,2),size(V,2)) )
      A%cV = V
      deflation_count = deflation_count + 1
      A%deflation_id = deflation_count
      do j = 1, size(V,2)
         A%cAV(:,j) = complex_mv( A, V(:,j) )
      end do
      A%cVtAV = matmul( conjg(transpose(V)), A%cAV )
   end subroutine cmake_deflation

! add_deflation (matrix_module.f90:1433-1623): x = V E^-1 V^H b +
! (I - V E^-1 V^H A) y; with shifts, E = V^H AV - (sigma - seed) I per shift
   function radd_deflation( A, b, y ) result(x)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: b(:), y(:)
      real(kind=rp)                     :: x(size(y))
      call idrs_check( idrs_add_deflation_f64(matrix_handle(A, .false.), b, y, x), 'add_deflation' )
   end function radd_deflation

   function cadd_deflation( A, b, y ) result(x)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: b(:), y(:)
      complex(kind=cp)                  :: x(size(y))
      call idrs_check( idrs_add_deflation_c128(matrix_handle(A, .true.), b, y, x), 'add_deflation' )
   end function cadd_deflation

   function rms_add_deflation( A, sigma, b, y ) result(x)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: sigma(:), b(:), y(:,:)
      real(kind=rp)                     :: x(size(y,1),size(y,2))
      call idrs_check( idrs_ms_add_deflation_f64(matrix_handle(A, .false.), size(sigma), sigma, b, y, x), &
                       'add_deflation' )
   end function rms_add_deflation

   function cms_add_deflation( A, sigma, b, y ) result(x)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: sigma(:), b(:), y(:,:)
      complex(kind=cp)                  :: x(size(y,1),size(y,2))
      call idrs_check( idrs_ms_add_deflation_c128(matrix_handle(A, .true.), size(sigma), sigma, b, y, x), &
                       'add_deflation' )
   end function cms_add_deflation

   subroutine real_preconditioner( A, preconditioner, degree, foci, center, H )
      type(matrix), intent(inout)                 :: A
      integer, intent(in), optional               :: preconditioner, degree
      real(kind=rp), intent(in), optional         :: foci(2), center
      real(kind=rp), intent(in), optional         :: H(:,:)
      real(kind=rp)                               :: f(2), c

      A%preconditioner = 0
      if ( present(preconditioner) ) A%preconditioner = preconditioner
      A%degree = 0
      if ( present(degree) ) A%degree = degree
      if ( allocated(A%ralpha_p) ) deallocate(A%ralpha_p)

      select case(A%preconditioner)
         case(1)
            if ( present(center) ) then
               c = center - A%rseed
            elseif ( present(H) ) then
               call analyse_spectrum( H, center = c )
            else
               c = 1._rp
            end if
            allocate(A%ralpha_p(0:A%degree))
            A%ralpha_p = neumann( A%degree, c )
         case(2)
            if ( present(foci) ) then
               f = foci - A%rseed
            elseif ( present(H) ) then
               call analyse_spectrum( H, foci = f )
            else
               f(1) = 0.25
               f(2) = 1.75
            end if
            allocate(A%ralpha_p(0:A%degree))
            A%ralpha_p = chebyshev( A%degree, f )
         case(3)
            allocate(A%ralpha_p(0:A%degree))
            A%ralpha_p = compute_alpha( A%degree, H )
      end select
   end subroutine real_preconditioner

   subroutine complex_preconditioner( A, preconditioner, degree, foci, center, H )
      type(matrix), intent(inout)                 :: A
      integer, intent(in), optional               :: preconditioner, degree
      complex(kind=cp), intent(in), optional      :: foci(2), center
      complex(kind=cp), intent(in), optional      :: H(:,:)
      complex(kind=cp)                            :: f(2), c

      A%preconditioner = 0
      if ( present(preconditioner) ) A%preconditioner = preconditioner
      A%degree = 0
      if ( present(degree) ) A%degree = degree
      if ( allocated(A%calpha_p) ) deallocate(A%calpha_p)

      select case(A%preconditioner)
         case(1)
            if ( present(center) ) then
               c = center - A%cseed
            elseif ( present(H) ) then
               call analyse_spectrum( H, center = c )
            else
               c = 1._cp
            end if
            allocate(A%calpha_p(0:A%degree))
            A%calpha_p = neumann( A%degree, c )
         case(2)
            if ( present(foci) ) then
               f = foci - A%cseed
            elseif ( present(H) ) then
               call analyse_spectrum( H, foci = f )
            else
               f(1) = 0.25_cp
               f(2) = 1.75_cp
            end if
            allocate(A%calpha_p(0:A%degree))
            A%calpha_p = chebyshev( A%degree, f )
         case(3)
            allocate(A%calpha_p(0:A%degree))
            A%calpha_p = compute_alpha( A%degree, H )
      end select
   end subroutine complex_preconditioner

! filter_ritzval (matrix_module.f90:1059-1077)
   function filter_ritzval( ritzval ) result(outlier)
      complex(kind=cp), intent(in)      :: ritzval(:)
      logical                           :: outlier(size(ritzval))
      integer(c_int)                    :: o(size(ritzval))
      call idrs_check( idrs_filter_ritzval(size(ritzval), ritzval, o), 'filter_ritzval' )
      outlier = o /= 0
   end function filter_ritzval

! ranalyse_spectrum (matrix_module.f90:1079-1181)
   subroutine ranalyse_spectrum( H, foci, center, shift )
      real(kind=rp), intent(in)            :: H(:,:)
      real(kind=rp), intent(out), optional :: foci(2), center, shift
      complex(kind=cp), allocatable        :: Hc(:,:)
      complex(kind=cp)                     :: f(2), c, sh
      Hc = cmplx(H, 0._rp, kind=cp)
      call idrs_check( idrs_analyse_spectrum(size(H,2), Hc, size(H,1), 0, f, c, sh), 'analyse_spectrum' )
      if ( present(foci) )   foci = real(f, kind=rp)
      if ( present(center) ) center = real(c, kind=rp)
      if ( present(shift) )  shift = real(sh, kind=rp)
   end subroutine ranalyse_spectrum

! canalyse_spectrum (matrix_module.f90:1183-1282)
   subroutine canalyse_spectrum( H, foci, center, shift )
      complex(kind=cp), intent(in)            :: H(:,:)
      complex(kind=cp), intent(out), optional :: foci(2), center, shift
      complex(kind=cp)                        :: f(2), c, sh
      call idrs_check( idrs_analyse_spectrum(size(H,2), H, size(H,1), 1, f, c, sh), 'analyse_spectrum' )
      if ( present(foci) )   foci = f
      if ( present(center) ) center = c
      if ( present(shift) )  shift = sh
   end subroutine canalyse_spectrum

   function rchebyshev( degree, foci ) result(alpha)
      integer, intent(in)                      :: degree
      real(kind=rp)                            :: alpha(0:degree)
      real(kind=rp), intent(in)                :: foci(2)
      complex(kind=cp)                         :: f(2), a(0:degree)
      f = cmplx(foci, 0._rp, kind=cp)
      call idrs_check( idrs_chebyshev_coeffs(degree, f, 0, a), 'chebyshev' )
      alpha = real(a, kind=rp)
   end function rchebyshev

   function cchebyshev( degree, foci ) result(alpha)
      integer, intent(in)                      :: degree
      complex(kind=cp)                         :: alpha(0:degree)
      complex(kind=cp), intent(in)             :: foci(2)
      call idrs_check( idrs_chebyshev_coeffs(degree, foci, 1, alpha), 'chebyshev' )
   end function cchebyshev

   function rneumann( degree, center ) result(alpha)
      integer, intent(in)                      :: degree
      real(kind=rp)                            :: alpha(0:degree)
      real(kind=rp), intent(in)                :: center
      complex(kind=cp)                         :: a(0:degree)
      call idrs_check( idrs_neumann_coeffs(degree, cmplx(center, 0._rp, kind=cp), 0, a), 'neumann' )
      alpha = real(a, kind=rp)
   end function rneumann

   function cneumann( degree, center ) result(alpha)
      integer, intent(in)                      :: degree
      complex(kind=cp)                         :: alpha(0:degree)
      complex(kind=cp), intent(in)             :: center
      call idrs_check( idrs_neumann_coeffs(degree, center, 1, alpha), 'neumann' )
   end function cneumann

   function rcompute_alpha( degree, H ) result(alpha)
      integer, intent(in)                         :: degree
      real(kind=rp)                               :: alpha(0:degree)
      real(kind=rp)                               :: H(:,:)
      complex(kind=cp), allocatable               :: Hc(:,:)
      complex(kind=cp)                            :: a(0:degree)
      Hc = cmplx(H, 0._rp, kind=cp)
      call idrs_check( idrs_compute_alpha_coeffs(degree, Hc, size(H,1), size(H,2), 0, a), 'compute_alpha' )
      alpha = real(a, kind=rp)
   end function rcompute_alpha

   function ccompute_alpha( degree, H ) result(alpha)
      integer, intent(in)                         :: degree
      complex(kind=cp)                            :: alpha(0:degree)
      complex(kind=cp)                            :: H(:,:)
      call idrs_check( idrs_compute_alpha_coeffs(degree, H, size(H,1), size(H,2), 1, alpha), 'compute_alpha' )
   end function ccompute_alpha

   function rscaleback( A, b, y ) result(x)
      type(matrix), intent(in)            :: A
      real(kind=rp), intent(in)           :: y(:)
      real(kind=rp), intent(in)           :: b(:)
      real(kind=rp)                       :: x(size(y))
      call idrs_check( idrs_scaleback_b_f64(matrix_handle(A, .false.), b, y, x), 'scaleback' )
   end function rscaleback

   function cscaleback( A, b, y ) result(x)
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)           :: y(:)
      complex(kind=cp), intent(in)           :: b(:)
      complex(kind=cp)                       :: x(size(y))
      call idrs_check( idrs_scaleback_b_c128(matrix_handle(A, .true.), b, y, x), 'scaleback' )
   end function cscaleback

! Scale back the multishift solutions (matrix_module.f90:724-806): column k
! of x is p_k(A) y(:,k) with the shifted polynomial alpha_s(:,k) from
! MS_PARAMETERS (alpha = 1 without a polynomial preconditioner), then the
! deflation step for sigma_k, the shift-and-invert factor
! (seed - shift)/(sigma_k - shift), then /A.
   function ms_rscaleback( A, sigma, b, y ) result(x)
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)              :: y(:,:)
      real(kind=rp), intent(in)              :: b(:)
      real(kind=rp), intent(in)              :: sigma(:)
      real(kind=rp)                          :: x(size(y,1),size(y,2))
      complex(kind=cp), allocatable          :: coef(:,:)
      integer                                :: deg
      deg = 0
      if ( A%degree > 0 ) then
         if ( .not. allocated(A%ralpha_s) ) error stop 'First call subroutine MS_PARAMETERS'
         deg = A%degree
         allocate( coef(0:deg, size(sigma)) )
         coef = cmplx( A%ralpha_s(0:deg, 1:size(sigma)), 0._rp, kind=cp )
      else
         allocate( coef(0:0, size(sigma)) )
         coef = (1._cp, 0._cp)
      end if
      call idrs_check( idrs_ms_scaleback_b_f64(matrix_handle(A, .false.), size(sigma), sigma, deg, coef, b, y, x), &
                       'scaleback' )
   end function ms_rscaleback

   function ms_cscaleback( A, sigma, b, y ) result(x)
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)           :: y(:,:)
      complex(kind=cp), intent(in)           :: b(:)
      complex(kind=cp), intent(in)           :: sigma(:)
      complex(kind=cp)                       :: x(size(y,1),size(y,2))
      complex(kind=cp), allocatable          :: coef(:,:)
      integer                                :: deg
      deg = 0
      if ( A%degree > 0 ) then
         if ( .not. allocated(A%calpha_s) ) error stop 'First call subroutine MS_PARAMETERS'
         deg = A%degree
         allocate( coef(0:deg, size(sigma)) )
         coef = A%calpha_s(0:deg, 1:size(sigma))
      else
         allocate( coef(0:0, size(sigma)) )
         coef = (1._cp, 0._cp)
      end if
      call idrs_check( idrs_ms_scaleback_b_c128(matrix_handle(A, .true.), size(sigma), sigma, deg, coef, b, y, x), &
                       'scaleback' )
   end function ms_cscaleback

! Shifts of the preconditioned system and the scale-back polynomials
! (matrix_module.f90:402-549): host arithmetic on the shifts and the
! polynomial coefficients only.  With shift-and-invert the shifts are
! eta = (sigma - seed)/(sigma - shift); every shift is checked against the
! preconditioner's shift (the reference's loop bound there is read before
! it is set).
   subroutine real_ms_parameters( A, sigma, sigma_p )
      type(matrix), intent(inout)                 :: A
      real(kind=rp), intent(in)                   :: sigma(:)
      real(kind=rp), intent(out)                  :: sigma_p(:)
      real(kind=rp)                               :: sigma_s(size(sigma))
      integer                                     :: i, i_sigma, n_sigma
      if ( A%shift_invert ) then
         if ( A%rshift == A%rseed ) error stop &
            'Preconditioner shift should not be equal to seed, choose preconditioner shift and seed differently'
         do i_sigma = 1, size(sigma)
            if ( A%rshift == sigma(i_sigma) ) error stop &
               'Preconditioner shift should not be equal to one of the shifts, choose different preconditioner shift'
         end do
         sigma_s = ( sigma-A%rseed )/( sigma-A%rshift )
      else
         sigma_s = sigma - A%rseed
      end if
      if ( A%degree > 0 ) then
         if ( .not. allocated(A%ralpha_p) ) error stop 'First call subroutine real_preconditioner '
         n_sigma = size(sigma)
         if ( allocated(A%ralpha_s) ) deallocate(A%ralpha_s)
         allocate(A%ralpha_s(0:A%degree,n_sigma))
         A%ralpha_s = 0.
         do i_sigma = 1,n_sigma
            A%ralpha_s(A%degree,i_sigma) = A%ralpha_p(A%degree)
            do i = A%degree-1,0,-1
               A%ralpha_s(i,i_sigma) = A%ralpha_p(i) + sigma_s(i_sigma)*A%ralpha_s(i+1,i_sigma)
            end do
            sigma_p(i_sigma) = sigma_s(i_sigma)*A%ralpha_s(0,i_sigma)
         end do
      else
         sigma_p = sigma_s
      end if
   end subroutine real_ms_parameters

   subroutine complex_ms_parameters( A, sigma, sigma_p )
      type(matrix), intent(inout)                 :: A
      complex(kind=cp), intent(in)                :: sigma(:)
      complex(kind=cp), intent(out)               :: sigma_p(:)
      complex(kind=cp)                            :: sigma_s(size(sigma))
      integer                                     :: i, i_sigma, n_sigma
      if ( A%shift_invert ) then
         if ( A%cshift == A%cseed ) error stop &
            'Preconditioner shift should not be equal to seed, choose preconditioner shift and seed differently'
         do i_sigma = 1, size(sigma)
            if ( A%cshift == sigma(i_sigma) ) error stop &
               'Preconditioner shift should not be equal to one of the shifts, choose different preconditioner shift'
         end do
         sigma_s = ( sigma-A%cseed )/( sigma-A%cshift )
      else
         sigma_s = sigma - A%cseed
      end if
      if ( A%degree > 0 ) then
         if ( .not. allocated(A%calpha_p) ) error stop 'First call subroutine complex_preconditioner '
         n_sigma = size(sigma)
         if ( allocated(A%calpha_s) ) deallocate(A%calpha_s)
         allocate(A%calpha_s(0:A%degree,n_sigma))
         A%calpha_s = 0.
         do i_sigma = 1,n_sigma
            A%calpha_s(A%degree,i_sigma) = A%calpha_p(A%degree)
            do i = A%degree-1,0,-1
               A%calpha_s(i,i_sigma) = A%calpha_p(i) + sigma_s(i_sigma)*A%calpha_s(i+1,i_sigma)
            end do
            sigma_p(i_sigma) = sigma_s(i_sigma)*A%calpha_s(0,i_sigma)
         end do
      else
         sigma_p = sigma_s
      end if
   end subroutine complex_ms_parameters

end module matrix_module
