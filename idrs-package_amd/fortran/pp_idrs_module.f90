!
! pp_idrs_module: drop-in for the reference's MODULES/PP_IDRS/pp_idrs_module.f90
! (with its include files initialize.inc, real_matrix.inc,
! complex_matrix.inc, real_preconditioner.inc, complex_preconditioner.inc,
! recycle.inc, run_idrs.inc, run_mr_idrs.inc, run_ms_idrs.inc,
! run_ms_mr_idrs.inc): x = pp_idrs( K, b[, sigma], M, C, P, center, foci,
! seed, shift, V ) for one or several right-hand sides and one or several
! shifts, real or complex.  Each call builds type(matrix) A from K, M, C, P
! and the seed and shift, sets up the polynomial preconditioner (from
! initial IDRS_RITZ iterations when its parameters are not given), the
! deflation space V and the recycle space, solves with the chosen method on
! Pmul(A, b) and scales back -- every product on the device operator.
!
! The reference reads its solver settings from the command line; here they
! are module settings with the reference's defaults (IDR(4), tol 1e-6 as a
! default-real literal, maxit 1000, no polynomial, 32 Ritz values, no
! recycling, inner QMRIDR settings in_s 1, in_tol 1e-1, in_it 0), changed by
! pp_idrs_settings.  Plot output (gnuplot) is not produced.
!
module pp_idrs_module

   use precision_module
   use user_module
   use matrix_module
   use idrs_module
   use ritz_module

   implicit none

   private
   public :: pp_idrs, pp_idrs_settings

   interface pp_idrs
      module procedure RPP_IDRS, CPP_IDRS, RPP_MR_IDRS, CPP_MR_IDRS, RPP_MS_IDRS, CPP_MS_IDRS, &
                       RPP_MS_MR_IDRS, CPP_MS_MR_IDRS
   end interface

! settings (initialize.inc's defaults)
   character(len=8), save :: opt_method = 'idrs'
   integer, save          :: opt_s = 4
   real(kind=rp), save    :: opt_tol = 1e-6
   integer, save          :: opt_maxit = 1000
   logical, save          :: opt_recycle = .false.
   integer, save          :: opt_n_omega = 0
   integer, save          :: opt_in_s = 1
   real(kind=rp), save    :: opt_in_tol = 1e-1
   integer, save          :: opt_in_it = 0
   integer, save          :: opt_precon = 0     ! 1 Neumann, 2 Chebyshev, 3 Hessenberg
   integer, save          :: opt_degree = 0
   integer, save          :: opt_ritz = 32
   logical, save          :: opt_verbose = .true.

contains

! The command-line options of initialize.inc as arguments: method 'idrs',
! 'bicgstab', 'qmridr' or 'minsync' (s for idrs / qmridr), tolerance,
! maximum_iterations, recycle, n_omega, inner_s / inner_tolerance /
! inner_iterations (QMRIDR's inner IDRS), one of chebyshev / neumann /
! hessenberg = degree, ritz; verbose prints the reference's result lines.
   subroutine pp_idrs_settings( method, s, tolerance, maximum_iterations, recycle, n_omega, inner_s, &
                                inner_tolerance, inner_iterations, chebyshev, neumann, hessenberg, ritz, verbose )
      character(len=*), intent(in), optional :: method
      integer, intent(in), optional          :: s, maximum_iterations, n_omega, inner_s, inner_iterations
      integer, intent(in), optional          :: chebyshev, neumann, hessenberg, ritz
      real(kind=rp), intent(in), optional    :: tolerance, inner_tolerance
      logical, intent(in), optional          :: recycle, verbose
      if ( present(method) ) then
         if ( method /= 'idrs' .and. method /= 'bicgstab' .and. method /= 'qmridr' .and. method /= 'minsync' ) &
            error stop 'pp_idrs_settings: method must be idrs, bicgstab, qmridr or minsync'
         opt_method = method
      end if
      if ( present(s) ) opt_s = s
      if ( present(tolerance) ) opt_tol = tolerance
      if ( present(maximum_iterations) ) opt_maxit = maximum_iterations
      if ( present(recycle) ) opt_recycle = recycle
      if ( present(n_omega) ) opt_n_omega = n_omega
      if ( present(inner_s) ) opt_in_s = inner_s
      if ( present(inner_tolerance) ) opt_in_tol = inner_tolerance
      if ( present(inner_iterations) ) opt_in_it = inner_iterations
      if ( present(chebyshev) ) then
         opt_precon = 2
         opt_degree = chebyshev
      elseif ( present(neumann) ) then
         opt_precon = 1
         opt_degree = neumann
      elseif ( present(hessenberg) ) then
         opt_precon = 3
         opt_degree = hessenberg
      end if
      if ( opt_degree == 0 ) opt_precon = 0
      if ( present(ritz) ) opt_ritz = ritz
      if ( present(verbose) ) opt_verbose = verbose
   end subroutine pp_idrs_settings

! s and IDRS variant of the method (initialize.inc: bicgstab is IDR(1)
! variant 2, minsync variant 3 with s = the image count, qmridr variant 0)
   subroutine method_s( s, variant )
      integer, intent(out) :: s, variant
      select case ( opt_method )
      case ( 'bicgstab' )
         s = 1
         variant = 2
      case ( 'minsync' )
         s = 1
         variant = 3
      case ( 'qmridr' )
         s = opt_s
         variant = 0
      case default
         s = opt_s
         variant = 1
      end select
   end subroutine method_s

! real_matrix.inc / complex_matrix.inc
   subroutine r_setup( A, K, M, C, P, seed, shift, multishift )
      type(matrix), intent(inout)              :: A
      type(user_matrix), intent(in)            :: K
      type(user_matrix), intent(in), optional  :: M, C, P
      real(kind=rp), intent(in), optional      :: seed, shift
      logical, intent(in)                      :: multishift
      A%K = K
      if ( present(M) ) then
         A%mass_matrix = .true.
         A%M = M
      end if
      if ( present(C) ) then
         A%damping_matrix = .true.
         A%C = C
      end if
      if ( present(P) ) then
         A%P = P
         A%shift_invert = present(shift)
      end if
      A%multishift = multishift
      A%block_system = multishift .and. present(C)
      A%rseed = 0._rp
      if ( present(seed) ) A%rseed = seed
      A%rshift = 0._rp
      if ( present(shift) ) A%rshift = shift
      A%preconditioner = 0
      A%degree = 0
   end subroutine r_setup

   subroutine c_setup( A, K, M, C, P, seed, shift, multishift )
      type(matrix), intent(inout)              :: A
      type(user_matrix), intent(in)            :: K
      type(user_matrix), intent(in), optional  :: M, C, P
      complex(kind=cp), intent(in), optional   :: seed, shift
      logical, intent(in)                      :: multishift
      A%K = K
      if ( present(M) ) then
         A%mass_matrix = .true.
         A%M = M
      end if
      if ( present(C) ) then
         A%damping_matrix = .true.
         A%C = C
      end if
      if ( present(P) ) then
         A%P = P
         A%shift_invert = present(shift)
      end if
      A%multishift = multishift
      A%block_system = multishift .and. present(C)
      A%cseed = (0._cp, 0._cp)
      if ( present(seed) ) A%cseed = seed
      A%cshift = (0._cp, 0._cp)
      if ( present(shift) ) A%cshift = shift
      A%preconditioner = 0
      A%degree = 0
   end subroutine c_setup

! real_preconditioner.inc: the polynomial's parameters are estimated from
! the Hessenberg matrix of initial IDRS_RITZ iterations unless given (a
! center for Neumann, foci for Chebyshev); the same iterations give the
! omegas when n_omega > 0.
   subroutine r_precondition( A, b, M1, s, center, foci, omega )
      type(matrix), intent(inout)              :: A
      real(kind=rp), intent(in)                :: b(:)
      type(matrix), intent(in)                 :: M1
      integer, intent(in)                      :: s
      real(kind=rp), intent(in), optional      :: center, foci(2)
      real(kind=rp), allocatable, intent(out)  :: omega(:)
      real(kind=rp), allocatable               :: H(:,:), x0(:)
      real(kind=rp)                            :: rc, rf(2), relres
      integer                                  :: n_ritz, ini_iter, flag, iter
      logical                                  :: estimate
      rc = 1._rp
      rf = [0.15_rp, 1.85_rp]
      if ( present(center) ) rc = center
      if ( present(foci) ) rf = foci
      estimate = .not. ( opt_precon == 0 .or. (opt_precon == 1 .and. present(center)) .or. &
                         (opt_precon == 2 .and. present(foci)) )
      n_ritz = opt_ritz
      if ( estimate .or. opt_n_omega > 0 ) then
         ini_iter = n_ritz + n_ritz/s
         allocate( omega(opt_n_omega), H(ini_iter+1,ini_iter), x0(size(b)) )
         x0 = idrs_ritz( A, b, s, M1, opt_tol, ini_iter, 1, flag, relres, iter, omega = omega, H = H )
         n_ritz = (iter/(s+1))*s + mod(iter,s+1)
      end if
      if ( estimate ) then
         call real_preconditioner( A, opt_precon, opt_degree, H = H(1:n_ritz,1:n_ritz) )
      else
         call real_preconditioner( A, opt_precon, opt_degree, rf, rc )
      end if
   end subroutine r_precondition

   subroutine c_precondition( A, b, M1, s, center, foci, omega )
      type(matrix), intent(inout)                 :: A
      complex(kind=cp), intent(in)                :: b(:)
      type(matrix), intent(in)                    :: M1
      integer, intent(in)                         :: s
      complex(kind=cp), intent(in), optional      :: center, foci(2)
      complex(kind=cp), allocatable, intent(out)  :: omega(:)
      complex(kind=cp), allocatable               :: H(:,:), x0(:)
      complex(kind=cp)                            :: cc, cf(2)
      real(kind=rp)                               :: relres
      integer                                     :: n_ritz, ini_iter, flag, iter
      logical                                     :: estimate
      cc = (1._cp, 0._cp)
      cf = [(0.15_cp, 0._cp), (1.85_cp, 0._cp)]
      if ( present(center) ) cc = center
      if ( present(foci) ) cf = foci
      estimate = .not. ( opt_precon == 0 .or. (opt_precon == 1 .and. present(center)) .or. &
                         (opt_precon == 2 .and. present(foci)) )
      n_ritz = opt_ritz
      if ( estimate .or. opt_n_omega > 0 ) then
         ini_iter = n_ritz + n_ritz/s
         allocate( omega(opt_n_omega), H(ini_iter+1,ini_iter), x0(size(b)) )
         x0 = idrs_ritz( A, b, s, M1, opt_tol, ini_iter, 1, flag, relres, iter, omega = omega, H = H )
         n_ritz = (iter/(s+1))*s + mod(iter,s+1)
      end if
      if ( estimate ) then
         call complex_preconditioner( A, opt_precon, opt_degree, H = H(1:n_ritz,1:n_ritz) )
      else
         call complex_preconditioner( A, opt_precon, opt_degree, cf, cc )
      end if
   end subroutine c_precondition

   subroutine report( ms, iter, flag, relres, orig )
      real(kind=rp), intent(in) :: ms
      integer, intent(in)       :: iter, flag
      real(kind=rp), intent(in) :: relres(:), orig(:)
      if ( .not. opt_verbose ) return
      write(*,'(a,f8.2,a)') 'Solution time          = ', ms/1000._rp, 's.'
      write(*,'(a,i5)')     'Number of iterations   = ', iter
      write(*,'(a,*(e9.2))') 'Relative residual norm = ', relres
      if ( flag == 1 ) write(*,'(a)') 'Maximum number of iterations reached!'
      if ( flag == 2 ) write(*,'(a)') 'Accuracy above prescribed tolerance!'
      if ( flag == 3 ) write(*,'(a)') 'Break down!'
      write(*,'(a,*(e9.2))') 'Relative residual norm original system = ', orig
      write(*,*)
   end subroutine report

   real(kind=rp) function wall_ms()
      integer(kind=8) :: t, rate
      call system_clock( t, rate )
      wall_ms = 1000._rp*real(t, rp)/real(rate, rp)
   end function wall_ms

! the relative residual of the original system for x (run_idrs.inc:47-57)
   real(kind=rp) function r_orig( K, b, x, M, C, seed )
      type(user_matrix), intent(in)            :: K
      real(kind=rp), intent(in)                :: b(:), x(:)
      type(user_matrix), intent(in), optional  :: M, C
      real(kind=rp), intent(in), optional      :: seed
      if ( .not. present(seed) ) then
         r_orig = norm( b - K*x )/norm( b )
      elseif ( present(M) .and. present(C) ) then
         r_orig = norm( b - K*x - seed*(C*x) - seed**2*(M*x) )/norm( b )
      elseif ( present(C) ) then
         r_orig = norm( b - K*x - seed*(C*x) - seed**2*x )/norm( b )
      elseif ( present(M) ) then
         r_orig = norm( b - K*x + seed*(M*x) )/norm( b )
      else
         r_orig = norm( b - K*x + seed*x )/norm( b )
      end if
   end function r_orig

   real(kind=rp) function c_orig( K, b, x, M, C, seed )
      type(user_matrix), intent(in)            :: K
      complex(kind=cp), intent(in)             :: b(:), x(:)
      type(user_matrix), intent(in), optional  :: M, C
      complex(kind=cp), intent(in), optional   :: seed
      if ( .not. present(seed) ) then
         c_orig = norm( b - K*x )/norm( b )
      elseif ( present(M) .and. present(C) ) then
         c_orig = norm( b - K*x - seed*(C*x) - seed**2*(M*x) )/norm( b )
      elseif ( present(C) ) then
         c_orig = norm( b - K*x - seed*(C*x) - seed**2*x )/norm( b )
      elseif ( present(M) ) then
         c_orig = norm( b - K*x + seed*(M*x) )/norm( b )
      else
         c_orig = norm( b - K*x + seed*x )/norm( b )
      end if
   end function c_orig

! one solve of run_idrs.inc / run_mr_idrs.inc: IDRS (or QMRIDR) on bp from
! x0, then x0 = the preconditioned solution and x = scaleback(A, b, x0)
   subroutine r_solve( A, M1, s, variant, b, bp, x0, U0, omega, x, flag, relres, iter )
      type(matrix), intent(in)                 :: A, M1
      integer, intent(in)                      :: s, variant
      real(kind=rp), intent(in)                :: b(:), bp(:)
      real(kind=rp), intent(inout)             :: x0(:)
      real(kind=rp), allocatable, intent(in)   :: U0(:,:), omega(:)
      real(kind=rp), intent(out)               :: x(:)
      integer, intent(out)                     :: flag, iter
      real(kind=rp), intent(out)               :: relres
      if ( variant == 0 ) then
         x = qmridr( A, bp, s, M1, opt_tol, opt_maxit, flag, relres, iter, opt_in_s, opt_in_tol, opt_in_it, x0, omega )
      else
         x = idrs( A, bp, s, M1, opt_tol, opt_maxit, variant, flag, relres, iter, x0, U0, omega )
      end if
      x0 = x
      x = scaleback( A, b, x0 )
   end subroutine r_solve

   subroutine c_solve( A, M1, s, variant, b, bp, x0, U0, omega, x, flag, relres, iter )
      type(matrix), intent(in)                   :: A, M1
      integer, intent(in)                        :: s, variant
      complex(kind=cp), intent(in)               :: b(:), bp(:)
      complex(kind=cp), intent(inout)            :: x0(:)
      complex(kind=cp), allocatable, intent(in)  :: U0(:,:), omega(:)
      complex(kind=cp), intent(out)              :: x(:)
      integer, intent(out)                       :: flag, iter
      real(kind=rp), intent(out)                 :: relres
      if ( variant == 0 ) then
         x = qmridr( A, bp, s, M1, opt_tol, opt_maxit, flag, relres, iter, opt_in_s, opt_in_tol, opt_in_it, x0, omega )
      else
         x = idrs( A, bp, s, M1, opt_tol, opt_maxit, variant, flag, relres, iter, x0, U0, omega )
      end if
      x0 = x
      x = scaleback( A, b, x0 )
   end subroutine c_solve

! recycle.inc: an IDRS_RITZ solve that returns the recycle space U0(n, s)
   subroutine r_recycle( A, M1, s, b, U0 )
      type(matrix), intent(in)                 :: A, M1
      integer, intent(in)                      :: s
      real(kind=rp), intent(in)                :: b(:)
      real(kind=rp), allocatable, intent(out)  :: U0(:,:)
      real(kind=rp), allocatable               :: x0(:)
      real(kind=rp)                            :: relres
      integer                                  :: flag, iter
      if ( .not. opt_recycle ) return
      allocate( U0(size(b),s), x0(size(b)) )
      x0 = idrs_ritz( A, b, s, M1, tolerance = opt_tol, maximum_iterations = opt_maxit, variant = 1, &
                      flag = flag, relres = relres, iterations = iter, U0 = U0 )
   end subroutine r_recycle

   subroutine c_recycle( A, M1, s, b, U0 )
      type(matrix), intent(in)                    :: A, M1
      integer, intent(in)                         :: s
      complex(kind=cp), intent(in)                :: b(:)
      complex(kind=cp), allocatable, intent(out)  :: U0(:,:)
      complex(kind=cp), allocatable               :: x0(:)
      real(kind=rp)                               :: relres
      integer                                     :: flag, iter
      if ( .not. opt_recycle ) return
      allocate( U0(size(b),s), x0(size(b)) )
      x0 = idrs_ritz( A, b, s, M1, tolerance = opt_tol, maximum_iterations = opt_maxit, variant = 1, &
                      flag = flag, relres = relres, iterations = iter, U0 = U0 )
   end subroutine c_recycle

! ---------------------------------------------------------------- one rhs
   function RPP_IDRS( K, b, M, C, P, center, foci, seed, shift, V ) result(x)
      type(user_matrix), intent(in)            :: K
      real(kind=rp), intent(in)                :: b(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      real(kind=rp), intent(in), optional      :: center, foci(2), seed, shift, V(:,:)
      real(kind=rp)                            :: x(size(b))
      type(matrix)                             :: A, M1
      real(kind=rp), allocatable               :: x0(:), U0(:,:), omega(:)
      real(kind=rp)                            :: relres, t0
      integer                                  :: s, variant, flag, iter
      call method_s( s, variant )
      allocate( x0(size(b)) )
      x0 = 0._rp
      call r_setup( A, K, M, C, P, seed, shift, .false. )
      call r_precondition( A, b, M1, s, center, foci, omega )
      if ( present(V) ) call make_deflation( A, V )
      call r_recycle( A, M1, s, b, U0 )
      t0 = wall_ms()
      call r_solve( A, M1, s, variant, b, Pmul(A,b), x0, U0, omega, x, flag, relres, iter )
      call report( wall_ms() - t0, iter, flag, [relres], [r_orig( K, b, x, M, C, seed )] )
   end function RPP_IDRS

   function CPP_IDRS( K, b, M, C, P, center, foci, seed, shift, V ) result(x)
      type(user_matrix), intent(in)            :: K
      complex(kind=cp), intent(in)             :: b(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      complex(kind=cp), intent(in), optional   :: center, foci(2), seed, shift, V(:,:)
      complex(kind=cp)                         :: x(size(b))
      type(matrix)                             :: A, M1
      complex(kind=cp), allocatable            :: x0(:), U0(:,:), omega(:)
      real(kind=rp)                            :: relres, t0
      integer                                  :: s, variant, flag, iter
      call method_s( s, variant )
      allocate( x0(size(b)) )
      x0 = (0._cp, 0._cp)
      call c_setup( A, K, M, C, P, seed, shift, .false. )
      call c_precondition( A, b, M1, s, center, foci, omega )
      if ( present(V) ) call make_deflation( A, V )
      call c_recycle( A, M1, s, b, U0 )
      t0 = wall_ms()
      call c_solve( A, M1, s, variant, b, Pmul(A,b), x0, U0, omega, x, flag, relres, iter )
      call report( wall_ms() - t0, iter, flag, [relres], [c_orig( K, b, x, M, C, seed )] )
   end function CPP_IDRS

! -------------------------------------------- several right-hand sides
! run_mr_idrs.inc: the columns one after another, each started from the
! previous column's preconditioned solution
   function RPP_MR_IDRS( K, b_all, M, C, P, center, foci, seed, shift, V ) result(x_all)
      type(user_matrix), intent(in)            :: K
      real(kind=rp), intent(in)                :: b_all(:,:)
      type(user_matrix), intent(in), optional  :: M, C, P
      real(kind=rp), intent(in), optional      :: center, foci(2), seed, shift, V(:,:)
      real(kind=rp)                            :: x_all(size(b_all,1),size(b_all,2))
      type(matrix)                             :: A, M1
      real(kind=rp), allocatable               :: x0(:), x(:), U0(:,:), omega(:)
      real(kind=rp)                            :: relres, t0
      integer                                  :: s, variant, flag, iter, irhs
      call method_s( s, variant )
      allocate( x0(size(b_all,1)), x(size(b_all,1)) )
      x0 = 0._rp
      call r_setup( A, K, M, C, P, seed, shift, .false. )
      if ( present(V) ) call make_deflation( A, V )
      call r_precondition( A, b_all(:,1), M1, s, center, foci, omega )
      call r_recycle( A, M1, s, b_all(:,1), U0 )
      do irhs = 1, size(b_all,2)
         if ( opt_verbose ) write(*,'(a,i2)') 'Results for rhs ', irhs
         t0 = wall_ms()
         call r_solve( A, M1, s, variant, b_all(:,irhs), Pmul(A,b_all(:,irhs)), x0, U0, omega, x, flag, relres, iter )
         x_all(:,irhs) = x
         call report( wall_ms() - t0, iter, flag, [relres], [r_orig( K, b_all(:,irhs), x, M, C, seed )] )
      end do
   end function RPP_MR_IDRS

   function CPP_MR_IDRS( K, b_all, M, C, P, center, foci, seed, shift, V ) result(x_all)
      type(user_matrix), intent(in)            :: K
      complex(kind=cp), intent(in)             :: b_all(:,:)
      type(user_matrix), intent(in), optional  :: M, C, P
      complex(kind=cp), intent(in), optional   :: center, foci(2), seed, shift, V(:,:)
      complex(kind=cp)                         :: x_all(size(b_all,1),size(b_all,2))
      type(matrix)                             :: A, M1
      complex(kind=cp), allocatable            :: x0(:), x(:), U0(:,:), omega(:)
      real(kind=rp)                            :: relres, t0
      integer                                  :: s, variant, flag, iter, irhs
      call method_s( s, variant )
      allocate( x0(size(b_all,1)), x(size(b_all,1)) )
      x0 = (0._cp, 0._cp)
      call c_setup( A, K, M, C, P, seed, shift, .false. )
      if ( present(V) ) call make_deflation( A, V )
      call c_precondition( A, b_all(:,1), M1, s, center, foci, omega )
      call c_recycle( A, M1, s, b_all(:,1), U0 )
      do irhs = 1, size(b_all,2)
         if ( opt_verbose ) write(*,'(a,i2)') 'Results for rhs ', irhs
         t0 = wall_ms()
         call c_solve( A, M1, s, variant, b_all(:,irhs), Pmul(A,b_all(:,irhs)), x0, U0, omega, x, flag, relres, iter )
         x_all(:,irhs) = x
         call report( wall_ms() - t0, iter, flag, [relres], [c_orig( K, b_all(:,irhs), x, M, C, seed )] )
      end do
   end function CPP_MR_IDRS

! ------------------------------------------------------------- multishift
! run_ms_idrs.inc: ms_parameters, MSIDRS (or MSQMRIDR) on Pmul(A, b) with
! the preconditioned shifts, the per-shift scale-back; with a damping
! matrix the block system of 2n rows ([-rhs; 0], the second half returned)
   subroutine r_ms_solve( A, M1, s, variant, sigma, b, X, flag, relres, iter, omega )
      type(matrix), intent(inout)              :: A
      type(matrix), intent(in)                 :: M1
      integer, intent(in)                      :: s, variant
      real(kind=rp), intent(in)                :: sigma(:), b(:)
      real(kind=rp), intent(out)               :: X(:,:), relres(:)
      integer, intent(out)                     :: flag, iter
      real(kind=rp), allocatable, intent(in)   :: omega(:)
      real(kind=rp)                            :: sigma_p(size(sigma))
      call ms_parameters( A, sigma, sigma_p )
      if ( variant == 0 ) then
         X = msqmridr( A, Pmul(A,b), sigma_p, s, M1, opt_tol, opt_maxit, flag, relres, iter, opt_in_s, opt_in_tol, &
                       opt_in_it, omega )
      else
         X = msidrs( A, Pmul(A,b), sigma_p, s, M1, opt_tol, opt_maxit, variant, flag, relres, iter, omega )
      end if
      X = scaleback( A, sigma, b, X )
   end subroutine r_ms_solve

   subroutine c_ms_solve( A, M1, s, variant, sigma, b, X, flag, relres, iter, omega )
      type(matrix), intent(inout)                 :: A
      type(matrix), intent(in)                    :: M1
      integer, intent(in)                         :: s, variant
      complex(kind=cp), intent(in)                :: sigma(:), b(:)
      complex(kind=cp), intent(out)               :: X(:,:)
      real(kind=rp), intent(out)                  :: relres(:)
      integer, intent(out)                        :: flag, iter
      complex(kind=cp), allocatable, intent(in)   :: omega(:)
      complex(kind=cp)                            :: sigma_p(size(sigma))
      call ms_parameters( A, sigma, sigma_p )
      if ( variant == 0 ) then
         X = msqmridr( A, Pmul(A,b), sigma_p, s, M1, opt_tol, opt_maxit, flag, relres, iter, opt_in_s, opt_in_tol, &
                       opt_in_it, omega )
      else
         X = msidrs( A, Pmul(A,b), sigma_p, s, M1, opt_tol, opt_maxit, variant, flag, relres, iter, omega )
      end if
      X = scaleback( A, sigma, b, X )
   end subroutine c_ms_solve

   function RPP_MS_IDRS( K, rhs, sigma, M, C, P, center, foci, seed, shift, V ) result(solution)
      type(user_matrix), intent(in)            :: K
      real(kind=rp), intent(in)                :: rhs(:), sigma(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      real(kind=rp), intent(in), optional      :: center, foci(2), seed, shift, V(:,:)
      real(kind=rp)                            :: solution(size(rhs),size(sigma))
      real(kind=rp), allocatable               :: solution_mr(:,:,:)
      solution_mr = RPP_MS_MR_IDRS( K, reshape(rhs, [size(rhs),1]), sigma, M, C, P, center, foci, seed, shift, V )
      solution = solution_mr(:,:,1)
   end function RPP_MS_IDRS

   function CPP_MS_IDRS( K, rhs, sigma, M, C, P, center, foci, seed, shift, V ) result(solution)
      type(user_matrix), intent(in)            :: K
      complex(kind=cp), intent(in)             :: rhs(:), sigma(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      complex(kind=cp), intent(in), optional   :: center, foci(2), seed, shift, V(:,:)
      complex(kind=cp)                         :: solution(size(rhs),size(sigma))
      complex(kind=cp), allocatable            :: solution_mr(:,:,:)
      solution_mr = CPP_MS_MR_IDRS( K, reshape(rhs, [size(rhs),1]), sigma, M, C, P, center, foci, seed, shift, V )
      solution = solution_mr(:,:,1)
   end function CPP_MS_IDRS

! run_ms_mr_idrs.inc (the single right-hand side is its first column: the
! reference's RPP_MS_IDRS / CPP_MS_IDRS run the same steps)
   function RPP_MS_MR_IDRS( K, rhs, sigma, M, C, P, center, foci, seed, shift, V ) result(solution)
      type(user_matrix), intent(in)            :: K
      real(kind=rp), intent(in)                :: rhs(:,:), sigma(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      real(kind=rp), intent(in), optional      :: center, foci(2), seed, shift, V(:,:)
      real(kind=rp)                            :: solution(size(rhs,1),size(sigma),size(rhs,2))
      type(matrix)                             :: A, M1
      real(kind=rp), allocatable               :: b_all(:,:), X(:,:), omega(:)
      real(kind=rp)                            :: relres(size(sigma)), orig(size(sigma)), t0
      integer                                  :: s, variant, flag, iter, nn, neq, irhs, i_sigma
      call method_s( s, variant )
      nn = size(rhs,1)
      neq = nn
      if ( present(C) ) neq = 2*nn
      allocate( b_all(neq,size(rhs,2)), X(neq,size(sigma)) )
      if ( present(C) ) then
         b_all(1:nn,:) = -rhs
         b_all(nn+1:,:) = 0._rp
      else
         b_all = rhs
      end if
      call r_setup( A, K, M, C, P, seed, shift, .true. )
      call r_precondition( A, b_all(:,1), M1, s, center, foci, omega )
      if ( present(V) ) call make_deflation( A, V )
      do irhs = 1, size(rhs,2)
         if ( opt_verbose .and. size(rhs,2) > 1 ) write(*,'(a,i2)') 'Results for rhs ', irhs
         t0 = wall_ms()
         call r_ms_solve( A, M1, s, variant, sigma, b_all(:,irhs), X, flag, relres, iter, omega )
         solution(:,:,irhs) = X(neq-nn+1:neq,:)
         do i_sigma = 1, size(sigma)
            orig(i_sigma) = r_orig( K, rhs(:,irhs), solution(:,i_sigma,irhs), M, C, sigma(i_sigma) )
         end do
         call report( wall_ms() - t0, iter, flag, relres, orig )
      end do
   end function RPP_MS_MR_IDRS

   function CPP_MS_MR_IDRS( K, rhs, sigma, M, C, P, center, foci, seed, shift, V ) result(solution)
      type(user_matrix), intent(in)            :: K
      complex(kind=cp), intent(in)             :: rhs(:,:), sigma(:)
      type(user_matrix), intent(in), optional  :: M, C, P
      complex(kind=cp), intent(in), optional   :: center, foci(2), seed, shift, V(:,:)
      complex(kind=cp)                         :: solution(size(rhs,1),size(sigma),size(rhs,2))
      type(matrix)                             :: A, M1
      complex(kind=cp), allocatable            :: b_all(:,:), X(:,:), omega(:)
      real(kind=rp)                            :: relres(size(sigma)), orig(size(sigma)), t0
      integer                                  :: s, variant, flag, iter, nn, neq, irhs, i_sigma
      call method_s( s, variant )
      nn = size(rhs,1)
      neq = nn
      if ( present(C) ) neq = 2*nn
      allocate( b_all(neq,size(rhs,2)), X(neq,size(sigma)) )
      if ( present(C) ) then
         b_all(1:nn,:) = -rhs
         b_all(nn+1:,:) = (0._cp, 0._cp)
      else
         b_all = rhs
      end if
      call c_setup( A, K, M, C, P, seed, shift, .true. )
      call c_precondition( A, b_all(:,1), M1, s, center, foci, omega )
      if ( present(V) ) call make_deflation( A, V )
      do irhs = 1, size(rhs,2)
         if ( opt_verbose .and. size(rhs,2) > 1 ) write(*,'(a,i2)') 'Results for rhs ', irhs
         t0 = wall_ms()
         call c_ms_solve( A, M1, s, variant, sigma, b_all(:,irhs), X, flag, relres, iter, omega )
         solution(:,:,irhs) = X(neq-nn+1:neq,:)
         do i_sigma = 1, size(sigma)
            orig(i_sigma) = c_orig( K, rhs(:,irhs), solution(:,i_sigma,irhs), M, C, sigma(i_sigma) )
         end do
         call report( wall_ms() - t0, iter, flag, relres, orig )
      end do
   end function CPP_MS_MR_IDRS

end module pp_idrs_module
