!
! mm_module: drop-in for the reference's EXAMPLES/MATRIX_MARKET/mm_module.f90
! (mm_matrix, :48-150) over libidrs_hip.so's Matrix Market reader.
!
!   A = mm_matrix( mm_file [, which_format] )
!
! A coordinate file becomes a user matrix stored on the device ('coo', the
! default, or 'crs'), with the symmetric, skew-symmetric and hermitian
! halves mirrored as mm_coo / mm_crs do; an array file becomes a dense user
! matrix holding the values on the host (rmat / cmat), as mm_dense does.
! One image: every row is local.
!
module mm_module

   use precision_module
   use idrs_capi
   use user_module
   use iso_c_binding

   implicit none

   private
   public :: mm_matrix

contains

   function mm_matrix( mm_file, which_format ) result(A)
      character(len=*), intent(in)           :: mm_file
      character(len=3), optional, intent(in) :: which_format
      type(user_matrix)                      :: A
      character(kind=c_char)                 :: path(len_trim(mm_file)+1)
      character(len=3)                       :: format
      integer(c_int)                         :: rep, field, symm, nrows, ncols
      integer(c_int64_t)                     :: entries
      real(kind=rp), allocatable             :: vals(:)
      integer                                :: i

      format = 'coo'
      if ( present(which_format) ) format = which_format
      do i = 1, len_trim(mm_file)
         path(i) = mm_file(i:i)
      end do
      path(len_trim(mm_file)+1) = c_null_char
      call idrs_check( idrs_mm_info(path, rep, field, symm, nrows, ncols, entries), 'mm_matrix' )
      if ( rep == 1 ) then
! coordinate: CRS on the device (3 'coo', 2 'crs')
         A%matrix_format = 3
         if ( format == 'crs' ) A%matrix_format = 2
         call idrs_check( idrs_mm_matrix(path, A%matrix_format, A%h), 'mm_matrix' )
         A%nrows = nrows
         A%ncols = ncols
         A%is_complex = field == 2
         call user_matrix_sizes( A )
      else
! array: the values on the host
         if ( field == 2 ) then
            allocate( vals(2*nrows*ncols) )
            call idrs_check( idrs_mm_dense(path, vals), 'mm_matrix' )
            A = dense_format( cmat = reshape( cmplx( vals(1::2), vals(2::2), kind=cp ), [nrows, ncols] ) )
         else
            allocate( vals(nrows*ncols) )
            call idrs_check( idrs_mm_dense(path, vals), 'mm_matrix' )
            A = dense_format( rmat = reshape( vals, [nrows, ncols] ) )
         end if
      end if
   end function mm_matrix

end module mm_module
