!
! Drop-in check: reference-style calls of the IDRS package interface,
! compiled against the GPU library's Fortran modules.  Prints one line per
! case: "CASE <name> iter <n> flag <f> relres <r> err <e>".
!
program drop_in

   use precision_module
   use user_module
   use mm_module
   use matrix_module
   use idrs_module
   use ritz_module
   use pp_idrs_module

   implicit none

   character(len=512) :: data_dir

   call toeplitz_real()
   call cdr_preconditioned()
   call toeplitz_complex()
   call krylov_family()
   call ms_pp_sequence()
   call cdr_ritz_chain()
! Matrix Market files: the directory is the first command-line argument
   if ( command_argument_count() >= 1 ) then
      call get_command_argument( 1, data_dir )
      call stommel_mm( trim(data_dir) )
      call wedge_neumann_deflated( trim(data_dir) )
      call wedge_sommerfeld_multi( trim(data_dir) )
      call wedge_neumann_lu_multi( trim(data_dir) )
      call wedge_sommerfeld_lu_multi( trim(data_dir) )
      call pp_idrs_chains( trim(data_dir) )
   end if

contains

! The stommel example's chain (EXAMPLES/MATRIX_MARKET/stommel.f90): the
! system matrix and the twelve monthly right-hand sides read by mm_matrix,
! diagonal scaling, here IDR(4) with the default Chebyshev(8) polynomial of
! real_preconditioner, for the first right-hand side; err is the true
! relative residual of the scaled-back solution.
   subroutine stommel_mm( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, RHS
      type(matrix)                  :: A, M1
      real(kind=rp), allocatable    :: x(:), b(:), y(:)
      real(kind=rp)                 :: relres
      integer                       :: flag, iter

      K = mm_matrix( dir//'/stommel6.mtx', 'coo' )
      RHS = mm_matrix( dir//'/stommel6_b.mtx' )
      allocate( b(K%nrows), x(K%nrows), y(K%nrows) )
      b = RHS%rmat(:,1)
      A%K = K
      A%P = real_precon( 1, K )
      call real_preconditioner( A, 2, 8 )
      y = idrs( A, b, 4, M1, 1.e-8_rp, 4000, 1, flag, relres, iter )
      x = scaleback( A, b, y )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'stommel6_mm', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', norm( b - K*x )/norm( b )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'stommel6_mm_sizes', ' iter', K%nnz, &
         ' flag', size(RHS%rmat,2), ' relres', real(K%nrows, rp), ' err', 0._rp
      call user_matrix_free( A%P )
      call user_matrix_free( K )
   end subroutine stommel_mm

! The wedge Neumann example's single-frequency chain with deflation
! (EXAMPLES/MATRIX_MARKET/wedge_neumann.f90 with PP_IDRS's RPP_IDRS steps
! written out): (K - sigma M) x = b, sigma = (2 pi)^2, V = M 1,
! make_deflation, IDR(4) on Pmul(A, b) at tol 1e-6 with the default M1,
! scaleback with the deflation step; err is the true relative residual.
   subroutine wedge_neumann_deflated( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, M, RHS
      type(matrix)                  :: A, M1
      real(kind=rp), allocatable    :: x(:), b(:), y(:), V(:,:)
      real(kind=rp)                 :: relres, sigma
      real(kind=rp), parameter      :: pi = 3.141592653589793_rp
      integer                       :: flag, iter, n

      K = mm_matrix( dir//'/wedge3_K.mtx' )
      M = mm_matrix( dir//'/wedge3_M.mtx' )
      RHS = mm_matrix( dir//'/wedge3_b.mtx' )
      n = K%nrows
      allocate( b(n), x(n), y(n), V(n,1) )
      b = RHS%rmat(:,1)
      sigma = (2._rp*pi)**2
      A%K = K
      A%mass_matrix = .true.
      A%M = M
      A%rseed = sigma
      V = 1._rp
      V(:,1) = M*V(:,1)
      call make_deflation( A, V )
      y = idrs( A, Pmul(A,b), 4, M1, 1.e-6_rp, 1000, 1, flag, relres, iter )
      x = scaleback( A, b, y )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'wedge3_neumann', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', norm( b - K*x + sigma*(M*x) )/norm( b )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'wedge3_neumann_vtav', ' iter', A%n_deflation, &
         ' flag', 0, ' relres', A%rVtAV(1,1), ' err', abs( A%rVtAV(1,1) - dot_product( V(:,1), K*V(:,1) - sigma*(M*V(:,1)) ) )
      call user_matrix_free( M )
      call user_matrix_free( K )
   end subroutine wedge_neumann_deflated

! The wedge Sommerfeld example with -multi -deflation (wedge_sommerfeld.f90,
! CPP_MS_IDRS of PP_IDRS written out): the block system of 2 n rows for
! K + sigma C + sigma^2 M at sigma = 2 pi f i, f = 1, 2; lumped mass
! preconditioner complex_precon(1, M); seed at the last frequency;
! V = [0; 1/sqrt(n)]; ms_parameters, MSIDRS(4) on Pmul(A, b), tol 1e-6;
! scaleback per shift; the solution is the second half of each column; err
! is each frequency's true relative residual of the damped system.
   subroutine wedge_sommerfeld_multi( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, M, C, RHS, P
      type(matrix)                  :: A, M1
      complex(kind=cp), allocatable :: x(:,:), b(:), rhs_c(:), V(:,:), sol(:,:)
      complex(kind=cp)              :: sigma(2), sigma_p(2)
      real(kind=rp)                 :: rr(2)
      real(kind=rp), parameter      :: pi = 3.141592653589793_rp
      integer                       :: flag, iter, nn, i
      character(len=24)             :: name

      K = mm_matrix( dir//'/wedge2_K.mtx' )
      M = mm_matrix( dir//'/wedge2_M.mtx' )
      C = mm_matrix( dir//'/wedge2_C.mtx' )
      RHS = mm_matrix( dir//'/wedge2_b.mtx' )
      nn = K%nrows
      allocate( b(2*nn), rhs_c(nn), x(2*nn,2), V(2*nn,1), sol(nn,2) )
      rhs_c = RHS%rmat(:,1)
      b(1:nn) = -rhs_c
      b(nn+1:) = (0._cp, 0._cp)
      sigma = (2._cp*pi*[1._cp, 2._cp])*(0._cp, 1._cp)
      P = complex_precon( 1, M )
      A%K = K
      A%mass_matrix = .true.
      A%M = M
      A%damping_matrix = .true.
      A%C = C
      A%P = P
      A%multishift = .true.
      A%block_system = .true.
      A%cseed = sigma(2)
      V = (0._cp, 0._cp)
      V(nn+1:,1) = 1._cp/sqrt(real(nn, kind=cp))
      call make_deflation( A, V )
      call ms_parameters( A, sigma, sigma_p )
      x = msidrs( A, Pmul(A,b), sigma_p, 4, M1, 1.e-6_rp, 1000, 1, flag, rr, iter )
      x = scaleback( A, sigma, b, x )
      sol = x(nn+1:,:)
      do i = 1, 2
         write(name,'(a,i0)') 'wedge2_sommerfeld_f', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( rhs_c - K*sol(:,i) - sigma(i)*(C*sol(:,i)) - sigma(i)**2*(M*sol(:,i)) ) &
            /norm( rhs_c )
      end do
      call user_matrix_free( P )
      call user_matrix_free( C )
      call user_matrix_free( M )
      call user_matrix_free( K )
   end subroutine wedge_sommerfeld_multi

! wedge_neumann.f90 -lu -multi (RPP_MS_IDRS written out): the band LU
! shift-and-invert preconditioner P = real_precon(3, K, M, shift) at
! shift = -2 s_1 s_3/(s_1 + s_3), seed s_3, frequencies 1, 2, 4;
! ms_parameters, MSIDRS(4) on Pmul(A, b) at tol 1e-6, scaleback; err is each
! frequency's true relative residual of K - sigma M.
   subroutine wedge_neumann_lu_multi( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, M, RHS, P
      type(matrix)                  :: A, M1
      real(kind=rp), allocatable    :: x(:,:), b(:)
      real(kind=rp)                 :: sigma(3), sigma_p(3), rr(3), rshift
      real(kind=rp), parameter      :: pi = 3.141592653589793_rp
      integer                       :: flag, iter, n, i
      character(len=24)             :: name

      K = mm_matrix( dir//'/wedge3_K.mtx' )
      M = mm_matrix( dir//'/wedge3_M.mtx' )
      RHS = mm_matrix( dir//'/wedge3_b.mtx' )
      n = K%nrows
      allocate( b(n), x(n,3) )
      b = RHS%rmat(:,1)
      sigma = (2._rp*pi*[1._rp, 2._rp, 4._rp])**2
      rshift = -2._rp*sigma(1)*sigma(3)/(sigma(1)+sigma(3))
      P = real_precon( 3, K, M, shift=rshift )
      A%K = K
      A%mass_matrix = .true.
      A%M = M
      A%P = P
      A%shift_invert = .true.
      A%rshift = rshift
      A%multishift = .true.
      A%rseed = sigma(3)
      call ms_parameters( A, sigma, sigma_p )
      x = msidrs( A, Pmul(A,b), sigma_p, 4, M1, 1.e-6_rp, 1000, 1, flag, rr, iter )
      x = scaleback( A, sigma, b, x )
      do i = 1, 3
         write(name,'(a,i0)') 'wedge3_neumann_lu_f', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( b - K*x(:,i) + sigma(i)*(M*x(:,i)) )/norm( b )
      end do
      call user_matrix_free( P )
      call user_matrix_free( M )
      call user_matrix_free( K )
   end subroutine wedge_neumann_lu_multi

! wedge_sommerfeld.f90 -lu -multi (CPP_MS_IDRS written out): the block
! system with the band LU of K + shift C + shift^2 M,
! shift = 2 pi/(f_3 + f_1) ((f_3 - f_1) sqrt(f_1 f_3) + 2 f_1 f_3 i),
! A%shift_invert, seed at the last of the frequencies 1, 2, 4; err is each
! frequency's true relative residual of the damped system.
   subroutine wedge_sommerfeld_lu_multi( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, M, C, RHS, P
      type(matrix)                  :: A, M1
      complex(kind=cp), allocatable :: x(:,:), b(:), rhs_c(:), sol(:,:)
      complex(kind=cp)              :: sigma(3), sigma_p(3), cshift
      real(kind=rp)                 :: rr(3), freq(3)
      real(kind=rp), parameter      :: pi = 3.141592653589793_rp
      integer                       :: flag, iter, nn, i
      character(len=24)             :: name

      K = mm_matrix( dir//'/wedge2_K.mtx' )
      M = mm_matrix( dir//'/wedge2_M.mtx' )
      C = mm_matrix( dir//'/wedge2_C.mtx' )
      RHS = mm_matrix( dir//'/wedge2_b.mtx' )
      nn = K%nrows
      allocate( b(2*nn), rhs_c(nn), x(2*nn,3), sol(nn,3) )
      rhs_c = RHS%rmat(:,1)
      b(1:nn) = -rhs_c
      b(nn+1:) = (0._cp, 0._cp)
      freq = [1._rp, 2._rp, 4._rp]
      sigma = (2._cp*pi*freq)*(0._cp, 1._cp)
      cshift = 2._cp*pi/(freq(3)+freq(1))*cmplx((freq(3)-freq(1))*sqrt(freq(1)*freq(3)), 2._rp*freq(1)*freq(3), &
               kind=cp)
      P = complex_precon( 3, K, M, C, shift=cshift )
      A%K = K
      A%mass_matrix = .true.
      A%M = M
      A%damping_matrix = .true.
      A%C = C
      A%P = P
      A%shift_invert = .true.
      A%cshift = cshift
      A%multishift = .true.
      A%block_system = .true.
      A%cseed = sigma(3)
      call ms_parameters( A, sigma, sigma_p )
      x = msidrs( A, Pmul(A,b), sigma_p, 4, M1, 1.e-6_rp, 1000, 1, flag, rr, iter )
      x = scaleback( A, sigma, b, x )
      sol = x(nn+1:,:)
      do i = 1, 3
         write(name,'(a,i0)') 'wedge2_sommerfeld_lu_f', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( rhs_c - K*sol(:,i) - sigma(i)*(C*sol(:,i)) - sigma(i)**2*(M*sol(:,i)) ) &
            /norm( rhs_c )
      end do
      call user_matrix_free( P )
      call user_matrix_free( C )
      call user_matrix_free( M )
      call user_matrix_free( K )
   end subroutine wedge_sommerfeld_lu_multi

! The same wedge chains through pp_idrs (PP_IDRS/pp_idrs_module.f90 with
! its default settings: IDR(4), tol 1e-6, maxit 1000): the Neumann single
! frequency with deflation (RPP_IDRS), the Sommerfeld multishift block
! system with the lumped mass preconditioner and deflation (CPP_MS_IDRS),
! and with the band LU shift-and-invert preconditioner (CPP_MS_IDRS with
! shift).  err: each frequency's true relative residual, recomputed here.
   subroutine pp_idrs_chains( dir )
      character(len=*), intent(in)  :: dir
      type(user_matrix)             :: K, M, C, RHS, P
      real(kind=rp), allocatable    :: b(:), x(:), V(:,:)
      complex(kind=cp), allocatable :: rhs_c(:), sol(:,:), Vc(:,:)
      complex(kind=cp)              :: sigma(3), cshift
      real(kind=rp)                 :: rsig, freq(3)
      real(kind=rp), parameter      :: pi = 3.141592653589793_rp
      integer                       :: n, i
      character(len=32)             :: name

      call pp_idrs_settings( verbose = .false. )
      K = mm_matrix( dir//'/wedge3_K.mtx' )
      M = mm_matrix( dir//'/wedge3_M.mtx' )
      RHS = mm_matrix( dir//'/wedge3_b.mtx' )
      n = K%nrows
      allocate( b(n), x(n), V(n,1) )
      b = RHS%rmat(:,1)
      rsig = (2._rp*pi)**2
      V = 1._rp
      V(:,1) = M*V(:,1)
      x = pp_idrs( K, b, M, P=P, seed=rsig, V=V )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'pp_wedge3_neumann', ' iter', 0, ' flag', 0, &
         ' relres', 0._rp, ' err', norm( b - K*x + rsig*(M*x) )/norm( b )
      call user_matrix_free( M )
      call user_matrix_free( K )

      K = mm_matrix( dir//'/wedge2_K.mtx' )
      M = mm_matrix( dir//'/wedge2_M.mtx' )
      C = mm_matrix( dir//'/wedge2_C.mtx' )
      RHS = mm_matrix( dir//'/wedge2_b.mtx' )
      n = K%nrows
      allocate( rhs_c(n), sol(n,3), Vc(2*n,1) )
      rhs_c = RHS%rmat(:,1)
      freq = [1._rp, 2._rp, 4._rp]
      sigma = (2._cp*pi*freq)*(0._cp, 1._cp)
      Vc = (0._cp, 0._cp)
      Vc(n+1:,1) = 1._cp/sqrt(real(n, kind=cp))
      P = complex_precon( 1, M )
      sol = pp_idrs( K, rhs_c, sigma, M, C, P, seed=sigma(3), V=Vc )
      do i = 1, 3
         write(name,'(a,i0)') 'pp_wedge2_multi_f', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', 0, ' flag', 0, &
            ' relres', 0._rp, ' err', norm( rhs_c - K*sol(:,i) - sigma(i)*(C*sol(:,i)) - sigma(i)**2*(M*sol(:,i)) ) &
            /norm( rhs_c )
      end do
      call user_matrix_free( P )
      cshift = 2._cp*pi/(freq(3)+freq(1))*cmplx((freq(3)-freq(1))*sqrt(freq(1)*freq(3)), 2._rp*freq(1)*freq(3), &
               kind=cp)
      P = complex_precon( 3, K, M, C, shift=cshift )
      sol = pp_idrs( K, rhs_c, sigma, M, C, P, seed=sigma(3), shift=cshift )
      do i = 1, 3
         write(name,'(a,i0)') 'pp_wedge2_lu_f', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', 0, ' flag', 0, &
            ' relres', 0._rp, ' err', norm( rhs_c - K*sol(:,i) - sigma(i)*(C*sol(:,i)) - sigma(i)**2*(M*sol(:,i)) ) &
            /norm( rhs_c )
      end do
      call user_matrix_free( P )
      call user_matrix_free( C )
      call user_matrix_free( M )
      call user_matrix_free( K )
   end subroutine pp_idrs_chains

! 1-D convection-diffusion, N=60 (EXAMPLES/TOEPLITZ/toeplitz.f90:111-165),
! stored in CRS; positional call as in the example.
   subroutine toeplitz_real()
      integer, parameter            :: N = 60, maxit = 4000
      real(kind=rp), parameter      :: Pe = 0.5
      type(matrix)                  :: A, M
      real(kind=rp)                 :: x(N), b(N), tol, relres
      integer                       :: row_ptr(N+1), col_ind(3*N), flag, iter, s, i, k, variant
      real(kind=rp)                 :: val(3*N)
      character(len=16)             :: name

      k = 0
      row_ptr(1) = 1
      do i = 1, N
         if ( i > 1 ) then
            k = k + 1; col_ind(k) = i-1; val(k) = -1._rp-Pe
         end if
         k = k + 1; col_ind(k) = i; val(k) = 2._rp
         if ( i < N ) then
            k = k + 1; col_ind(k) = i+1; val(k) = -1._rp+Pe
         end if
         row_ptr(i+1) = k + 1
      end do
      A%K = crs_format( N, N, row_ptr, col_ind(1:k), val(1:k) )
      b = 0.
      b(1) = 1._rp+Pe
      b(N) = 1._rp-Pe
      tol = 1.e-8_rp
      variant = 1
      do i = 1, 4
         s = 2**(i-1)
         x = idrs( A, b, s, M, tol, maxit, variant, flag, relres, iter )
         write(name,'(a,i0)') 'toeplitz_s', s
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', relres, ' err', maxval(abs(x-1._rp))
      end do
      x = idrs( A, b, 1, M, tol, maxit, 2, flag, relres, iter )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'toeplitz_bicgstab', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', maxval(abs(x-1._rp))
   end subroutine toeplitz_real

! 16^3 convection-diffusion-reaction matrix (Pe 5, reaction 1), 7-point
! stencil, CRS
   function cdr_matrix( m ) result(K)
      integer, intent(in)           :: m
      type(user_matrix)             :: K
      real(kind=rp), allocatable    :: val(:)
      integer, allocatable          :: row_ptr(:), col_ind(:)
      integer                       :: i, j, l, q, nz, n
      real(kind=rp)                 :: cw(7), h, e2, c
      integer                       :: off(7, 3)

      n = m*m*m
      h = 1._rp/(m+1)
      e2 = 1._rp/h**2
      c = 6._rp*e2 + 1._rp
      ! front, south, west, centre, east, north, back with Pe = 5
      cw = [ -e2-5._rp/(2*h), -e2-5._rp/(2*h), -e2-5._rp/(2*h), c, -e2+5._rp/(2*h), -e2+5._rp/(2*h), -e2+5._rp/(2*h) ]
      off(:,1) = [0, 0, -1, 0, 1, 0, 0]
      off(:,2) = [0, -1, 0, 0, 0, 1, 0]
      off(:,3) = [-1, 0, 0, 0, 0, 0, 1]
      allocate( row_ptr(n+1), col_ind(7*n), val(7*n) )
      nz = 0
      row_ptr(1) = 1
      do l = 0, m-1
         do j = 0, m-1
            do i = 0, m-1
               do q = 1, 7
                  if ( i+off(q,1) < 0 .or. i+off(q,1) >= m ) cycle
                  if ( j+off(q,2) < 0 .or. j+off(q,2) >= m ) cycle
                  if ( l+off(q,3) < 0 .or. l+off(q,3) >= m ) cycle
                  nz = nz + 1
                  col_ind(nz) = 1 + (i+off(q,1)) + m*(j+off(q,2)) + m*m*(l+off(q,3))
                  val(nz) = cw(q)
               end do
               row_ptr(1 + i + m*j + m*m*l + 1) = nz + 1
            end do
         end do
      end do
      K = crs_format( n, n, row_ptr, col_ind(1:nz), val(1:nz) )
   end function cdr_matrix

! pp_idrs call sequence (PP_IDRS/real_matrix.inc, real_preconditioner.inc,
! run_idrs.inc): diagonal scaling + Chebyshev(8), IDR(4), scale back.
   subroutine cdr_preconditioned()
      integer, parameter            :: m = 16, n = m*m*m
      type(matrix)                  :: A, M1
      type(user_matrix)             :: K
      real(kind=rp)                 :: x(n), b(n), x0(n), relres, foci(2)
      integer                       :: flag, iter

      K = cdr_matrix( m )
      x = 1._rp
      b = K*x
      A%K = K
      A%P = real_precon( 1, K )
      foci = [0.15_rp, 1.85_rp]
      call real_preconditioner( A, 2, 8, foci )
      x0 = idrs( A, Pmul(A,b), 4, M1, 1.e-8_rp, 2000, 1, flag, relres, iter )
      x = scaleback( A, b, x0 )
      relres = norm( b - K*x )/norm( b )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'cdr_cheb8', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', maxval(abs(x-1._rp))
      call user_matrix_free( A%P )
      call user_matrix_free( K )
   end subroutine cdr_preconditioned

! pp_idrs without foci (PP_IDRS/real_preconditioner.inc:12-34, run_idrs.inc):
! initial IDRS_RITZ iterations (n_ritz = 32, ini_iter = n_ritz + n_ritz/s)
! for the omegas and H, real_preconditioner(A, 2, 8, H = H(1:n_ritz,1:n_ritz))
! (foci from analyse_spectrum), then IDR(4) from the initial solution with
! those omegas, scale back.
   subroutine cdr_ritz_chain()
      integer, parameter            :: m = 16, n = m*m*m, s = 4, n_ritz = 32, ini = n_ritz + n_ritz/s
      type(matrix)                  :: A, M1
      type(user_matrix)             :: K
      real(kind=rp)                 :: x(n), b(n), x0(n), y(n), relres, omega(s), H(ini+1, ini)
      integer                       :: flag, iter, nr

      K = cdr_matrix( m )
      x = 1._rp
      b = K*x
      A%K = K
      A%P = real_precon( 1, K )
      x0 = idrs_ritz( A, b, s, M1, 1.e-8_rp, ini, 1, flag, relres, iter, omega = omega, H = H )
      nr = (iter/(s+1))*s + mod(iter,s+1)
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'cdr_ritz_init', ' iter', iter, ' flag', nr, &
         ' relres', relres, ' err', omega(1)
      call real_preconditioner( A, 2, 8, H = H(1:nr,1:nr) )
      y = idrs( A, Pmul(A,b), s, M1, 1.e-8_rp, 2000, 1, flag, relres, iter, x0, omega = omega )
      x = scaleback( A, b, y )
      relres = norm( b - K*x )/norm( b )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'cdr_ritz', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', maxval(abs(x-1._rp))
      call user_matrix_free( A%P )
      call user_matrix_free( K )
   end subroutine cdr_ritz_chain

! Complex Toeplitz accuracy test (toeplitz.f90:41-70) in offset storage.
   subroutine toeplitz_complex()
      integer, parameter            :: N = 200, maxit = 4000
      type(matrix)                  :: A, M
      complex(kind=cp)              :: x(N), b(N), vals(N,4), cv(4)
      integer                       :: offs(4), flag, iter, i, d, s, variant
      real(kind=rp)                 :: relres, tol
      character(len=32)             :: name

      offs = [-1, 0, 2, 3]
      cv = [ (0._cp,3.6_cp), (4._cp,0._cp), (1._cp,0._cp), (0.7_cp,0._cp) ]
      do d = 1, 4
         do i = 1, N
            if ( i+offs(d) >= 1 .and. i+offs(d) <= N ) then
               vals(i,d) = cv(d)
            else
               vals(i,d) = (0._cp,0._cp)
            end if
         end do
      end do
      A%K = dia_format( N, N, offs, cvals=vals )
      b = (0._cp,1._cp)
      ! the accuracy test's loops of toeplitz.f90:55-89: IDRS and QMRIDR
      ! with s = 1, 2, 4, ..., 64, BiCGSTAB between them
      tol = 1e-12_rp
      variant = 1
      do i = 1, 7
         s = 2**(i-1)
         x = idrs( A, b, s, M, tol, maxit, variant, flag, relres, iter )
         write(name,'(a,i0)') 'toeplitz_complex_idrs_s', s
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', relres, ' err', norm( b - A%K*x )/norm( b )
      end do
      x = idrs( A, b, 1, M, tol, maxit, 2, flag, relres, iter )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'toeplitz_complex_bicgstab', ' iter', iter, &
         ' flag', flag, ' relres', relres, ' err', norm( b - A%K*x )/norm( b )
      do i = 1, 7
         s = 2**(i-1)
         x = qmridr( A, b, s, M, tol, maxit, flag, relres, iter )
         write(name,'(a,i0)') 'toeplitz_complex_qmridr_s', s
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', relres, ' err', norm( b - A%K*x )/norm( b )
      end do
   end subroutine toeplitz_complex

! QMRIDR, MSIDRS and MSQMRIDR on the real Toeplitz system of toeplitz_real
! and its shifted versions (K - sigma I) x = b; err is the true relative
! residual of each shifted system.
   subroutine krylov_family()
      integer, parameter            :: N = 60, maxit = 4000, ns = 3
      real(kind=rp), parameter      :: Pe = 0.5
      type(matrix)                  :: A, M
      real(kind=rp)                 :: x(N), b(N), X2(N,ns), sig(ns), rr(ns), tol, relres
      integer                       :: row_ptr(N+1), col_ind(3*N), flag, iter, i, k
      real(kind=rp)                 :: val(3*N)
      character(len=24)             :: name

      k = 0
      row_ptr(1) = 1
      do i = 1, N
         if ( i > 1 ) then
            k = k + 1; col_ind(k) = i-1; val(k) = -1._rp-Pe
         end if
         k = k + 1; col_ind(k) = i; val(k) = 2._rp
         if ( i < N ) then
            k = k + 1; col_ind(k) = i+1; val(k) = -1._rp+Pe
         end if
         row_ptr(i+1) = k + 1
      end do
      A%K = crs_format( N, N, row_ptr, col_ind(1:k), val(1:k) )
      b = 0.
      b(1) = 1._rp+Pe
      b(N) = 1._rp-Pe
      tol = 1.e-8_rp
      x = qmridr( A, b, 4, M, tol, maxit, flag, relres, iter )
      write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', 'qmridr_s4', ' iter', iter, ' flag', flag, &
         ' relres', relres, ' err', maxval(abs(x-1._rp))
      sig = [0._rp, -0.5_rp, -2._rp]
      X2 = msidrs( A, b, sig, 4, tolerance=tol, maximum_iterations=maxit, flag=flag, relres=rr, iterations=iter )
      do i = 1, ns
         write(name,'(a,i0)') 'msidrs_shift', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( b - (A%K*X2(:,i) - sig(i)*X2(:,i)) )/norm( b )
      end do
      X2 = msqmridr( A, b, sig, 4, tolerance=tol, maximum_iterations=maxit, flag=flag, relres=rr, iterations=iter )
      do i = 1, ns
         write(name,'(a,i0)') 'msqmridr_shift', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( b - (A%K*X2(:,i) - sig(i)*X2(:,i)) )/norm( b )
      end do
   end subroutine krylov_family

! The pp_idrs multishift sequence (PP_IDRS/run_ms_idrs.inc): Chebyshev
! polynomial preconditioner, MS_PARAMETERS, MSIDRS on the preconditioned
! shifted systems, SCALEBACK per shift; err is the true relative residual of
! (K - sigma I) x = b for each shift (h^2-scaled 7-point Laplacian, m = 12).
   subroutine ms_pp_sequence()
      integer, parameter            :: m = 12, n = m*m*m, ns = 3
      type(matrix)                  :: A
      real(kind=rp)                 :: b(n), X(n,ns), sig(ns), sig_p(ns), rr(ns), foci(2), cw(7)
      integer                       :: row_ptr(n+1), col_ind(7*n), off(7,3), flag, iter, i, j, l, q, nz
      real(kind=rp)                 :: val(7*n)
      character(len=24)             :: name

      cw = [ -1._rp, -1._rp, -1._rp, 6._rp, -1._rp, -1._rp, -1._rp ]
      off(:,1) = [0, 0, -1, 0, 1, 0, 0]
      off(:,2) = [0, -1, 0, 0, 0, 1, 0]
      off(:,3) = [-1, 0, 0, 0, 0, 0, 1]
      nz = 0
      row_ptr(1) = 1
      do l = 0, m-1
         do j = 0, m-1
            do i = 0, m-1
               do q = 1, 7
                  if ( i+off(q,1) < 0 .or. i+off(q,1) >= m ) cycle
                  if ( j+off(q,2) < 0 .or. j+off(q,2) >= m ) cycle
                  if ( l+off(q,3) < 0 .or. l+off(q,3) >= m ) cycle
                  nz = nz + 1
                  col_ind(nz) = 1 + (i+off(q,1)) + m*(j+off(q,2)) + m*m*(l+off(q,3))
                  val(nz) = cw(q)
               end do
               row_ptr(1 + i + m*j + m*m*l + 1) = nz + 1
            end do
         end do
      end do
      A%K = crs_format( n, n, row_ptr, col_ind(1:nz), val(1:nz) )
      do i = 1, n
         b(i) = real(mod(i, 7), rp) - 3._rp
      end do
      foci = [0.05_rp, 11.95_rp]
      call real_preconditioner( A, 2, 4, foci )
      sig = [0._rp, -1._rp, -3._rp]
      call ms_parameters( A, sig, sig_p )
      X = msidrs( A, Pmul(A,b), sig_p, 4, tolerance=1.e-9_rp, maximum_iterations=2000, flag=flag, relres=rr, &
                  iterations=iter )
      X = scaleback( A, sig, b, X )
      do i = 1, ns
         write(name,'(a,i0)') 'ms_pp_shift', i
         write(*,'(a,1x,a,a,i6,a,i3,a,es12.4,a,es12.4)') 'CASE', trim(name), ' iter', iter, ' flag', flag, &
            ' relres', rr(i), ' err', norm( b - A%K*X(:,i) + sig(i)*X(:,i) )/norm( b )
      end do
   end subroutine ms_pp_sequence

end program drop_in
