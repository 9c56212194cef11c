!
! Working precision of the drop-in modules: the reference's
! MODULES/IDRS/precision_module.f90 kinds (real64 for real and complex).
!
module precision_module

   use iso_fortran_env, only: real64

   implicit none

   integer, parameter :: rp = real64   ! real precision
   integer, parameter :: cp = real64   ! complex precision

end module precision_module
