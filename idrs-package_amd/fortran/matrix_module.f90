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
!   Pmul          identity without deflation matrix_module.f90:1384-1431
!   filter_ritzval, ANALYSE_SPECTRUM         matrix_module.f90:1059-1282
!                 (host routines of the library)
!
! Not offered (stop with a message): deflation, mass/damping matrices,
! block systems, shift-and-invert.
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
             filter_ritzval, ANALYSE_SPECTRUM

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

! Cache of device operators built from type(matrix) states
   integer, parameter :: NCACHE = 8
   type :: op_entry
      type(c_ptr)                   :: op = c_null_ptr
      type(c_ptr)                   :: k = c_null_ptr
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

contains

! Device operator for the current state of A (cplx: complex arithmetic).
   function matrix_handle( A, cplx ) result(h)
      type(matrix), intent(in)          :: A
      logical, intent(in)               :: cplx
      type(c_ptr)                       :: h
      complex(kind=cp)                  :: seed, shift
      complex(kind=cp), allocatable     :: alpha(:)
      integer                           :: i, victim, deg
      integer(c_int)                    :: status

      if ( .not. c_associated(A%K%h) ) error stop 'matrix: the system matrix A%K is not set'
      if ( A%n_deflation > 0 ) error stop 'matrix: deflation is not offered by the GPU library'
      if ( A%block_system .or. A%shift_invert ) &
         error stop 'matrix: block systems and shift-and-invert are not offered by the GPU library'
      if ( (A%mass_matrix .or. A%damping_matrix) .and. (A%rseed /= 0._rp .or. A%cseed /= (0._cp,0._cp)) ) &
         error stop 'matrix: mass and damping matrices are not offered by the GPU library'
      if ( A%P%which_precon == 1 ) then
         if ( .not. c_associated(A%P%h, A%K%h) ) &
            error stop 'matrix: the diagonal preconditioner must be built from A%K'
      end if

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
      if ( A%P%which_precon == 1 ) then
         if ( A%P%has_shift ) then
            if ( cplx ) then
               call idrs_check( idrs_matrix_set_precon_c128(h, 1, shift), 'matrix' )
            else
               call idrs_check( idrs_matrix_set_precon_f64(h, 1, real(shift, kind=rp)), 'matrix' )
            end if
         else
            call idrs_check( idrs_matrix_set_precon_f64(h, 1), 'matrix' )
         end if
      end if
      call idrs_check( idrs_matrix_set_alpha(h, deg, alpha), 'matrix' )
      cache(victim)%op = h
      cache(victim)%k = A%K%h
      cache(victim)%which_precon = A%P%which_precon
      cache(victim)%has_shift = A%P%has_shift
      cache(victim)%shift = shift
      cache(victim)%seed = seed
      cache(victim)%cplx = cplx
      cache(victim)%degree = deg
      cache(victim)%alpha = alpha
      cache(victim)%stamp = clock
   end function matrix_handle

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
      if ( A%block_system ) error stop 'matrix: block systems are not offered by the GPU library'
      w = v/A%P
   end function rprecon

   function cprecon( v, A ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      if ( A%block_system ) error stop 'matrix: block systems are not offered by the GPU library'
      w = v/A%P
   end function cprecon

   function rPmul( A, v ) result(w)
      type(matrix), intent(in)          :: A
      real(kind=rp), intent(in)         :: v(:)
      real(kind=rp)                     :: w(size(v))
      if ( A%n_deflation > 0 ) error stop 'matrix: deflation is not offered by the GPU library'
      w = v
   end function rPmul

   function cPmul( A, v ) result(w)
      type(matrix), intent(in)          :: A
      complex(kind=cp), intent(in)      :: v(:)
      complex(kind=cp)                  :: w(size(v))
      if ( A%n_deflation > 0 ) error stop 'matrix: deflation is not offered by the GPU library'
      w = v
   end function cPmul

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
      call idrs_check( idrs_scaleback_f64(matrix_handle(A, .false.), y, x), 'scaleback' )
   end function rscaleback

   function cscaleback( A, b, y ) result(x)
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)           :: y(:)
      complex(kind=cp), intent(in)           :: b(:)
      complex(kind=cp)                       :: x(size(y))
      call idrs_check( idrs_scaleback_c128(matrix_handle(A, .true.), y, x), 'scaleback' )
   end function cscaleback

! Scale back the multishift solutions (matrix_module.f90:724-806, without
! deflation and shift-and-invert): column k of x is p_k(K') y(:,k) / D with
! the shifted polynomial alpha_s(:,k) from MS_PARAMETERS (alpha = 1 without
! a polynomial preconditioner).
   function ms_rscaleback( A, sigma, b, y ) result(x)
      type(matrix), intent(in)               :: A
      real(kind=rp), intent(in)              :: y(:,:)
      real(kind=rp), intent(in)              :: b(:)
      real(kind=rp), intent(in)              :: sigma(:)
      real(kind=rp)                          :: x(size(y,1),size(y,2))
      complex(kind=cp), allocatable          :: coef(:,:)
      integer                                :: deg
      if ( A%shift_invert ) error stop 'SCALEBACK: shift-and-invert is not offered by the GPU library'
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
      call idrs_check( idrs_ms_scaleback_f64(matrix_handle(A, .false.), size(sigma), deg, coef, y, x, 0), 'scaleback' )
   end function ms_rscaleback

   function ms_cscaleback( A, sigma, b, y ) result(x)
      type(matrix), intent(in)               :: A
      complex(kind=cp), intent(in)           :: y(:,:)
      complex(kind=cp), intent(in)           :: b(:)
      complex(kind=cp), intent(in)           :: sigma(:)
      complex(kind=cp)                       :: x(size(y,1),size(y,2))
      complex(kind=cp), allocatable          :: coef(:,:)
      integer                                :: deg
      if ( A%shift_invert ) error stop 'SCALEBACK: shift-and-invert is not offered by the GPU library'
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
      call idrs_check( idrs_ms_scaleback_c128(matrix_handle(A, .true.), size(sigma), deg, coef, y, x, 0), 'scaleback' )
   end function ms_cscaleback

! Shifts of the preconditioned system and the scale-back polynomials
! (matrix_module.f90:402-549, without shift-and-invert): host arithmetic on
! the polynomial coefficients only.
   subroutine real_ms_parameters( A, sigma, sigma_p )
      type(matrix), intent(inout)                 :: A
      real(kind=rp), intent(in)                   :: sigma(:)
      real(kind=rp), intent(out)                  :: sigma_p(:)
      real(kind=rp)                               :: sigma_s(size(sigma))
      integer                                     :: i, i_sigma, n_sigma
      if ( A%shift_invert ) error stop 'MS_PARAMETERS: shift-and-invert is not offered by the GPU library'
      sigma_s = sigma - A%rseed
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
      if ( A%shift_invert ) error stop 'MS_PARAMETERS: shift-and-invert is not offered by the GPU library'
      sigma_s = sigma - A%cseed
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
