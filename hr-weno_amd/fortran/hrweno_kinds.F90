module hrweno_kinds
!! Real kind of the library (src/hrweno_kinds.F90 of the reference).  The device kernels are
!! written for FP64 only, so rk is real64 whatever REAL32/REAL128 the caller defines.
   use, intrinsic :: iso_fortran_env, only: real64
   implicit none
   private

   public :: rk

   integer, parameter :: rk = real64

end module hrweno_kinds
