module hrw_check
!! Minimal checker for the Fortran test programs: counts failures, prints them, and stops
!! with a non-zero code at the end if any check failed.
   use hrweno_kinds, only: rk
   implicit none
   private
   public :: check, check_array, check_int, finish

   integer, save :: nfail = 0, npass = 0

contains

   subroutine check(name, actual, expected, thr, rel)
      character(*), intent(in) :: name
      real(rk), intent(in) :: actual, expected
      real(rk), intent(in), optional :: thr
      logical, intent(in), optional :: rel
      real(rk) :: tol, d
      logical :: r
      tol = epsilon(1.0_rk)
      if (present(thr)) tol = thr
      r = .false.
      if (present(rel)) r = rel
      d = abs(actual - expected)
      if (r) d = d/abs(expected)
      if (d > tol .or. d /= d) then
         nfail = nfail + 1
         write (*, '(a, a, 3es25.16)') "FAIL ", name, actual, expected, d
      else
         npass = npass + 1
      end if
   end subroutine check

   subroutine check_array(name, actual, expected, thr, rel)
      character(*), intent(in) :: name
      real(rk), intent(in) :: actual(:), expected(:)
      real(rk), intent(in), optional :: thr
      logical, intent(in), optional :: rel
      real(rk) :: tol, d
      logical :: r
      integer :: i
      tol = epsilon(1.0_rk)
      if (present(thr)) tol = thr
      r = .false.
      if (present(rel)) r = rel
      if (size(actual) /= size(expected)) then
         nfail = nfail + 1
         write (*, '(a, a, 2i10)') "FAIL size ", name, size(actual), size(expected)
         return
      end if
      do i = 1, size(actual)
         d = abs(actual(i) - expected(i))
         if (r) d = d/abs(expected(i))
         if (d > tol .or. d /= d) then
            nfail = nfail + 1
            write (*, '(a, a, i8, 3es25.16)') "FAIL ", name, i, actual(i), expected(i), d
            return
         end if
      end do
      npass = npass + 1
   end subroutine check_array

   subroutine check_int(name, actual, expected)
      character(*), intent(in) :: name
      integer, intent(in) :: actual, expected
      if (actual /= expected) then
         nfail = nfail + 1
         write (*, '(a, a, 2i12)') "FAIL ", name, actual, expected
      else
         npass = npass + 1
      end if
   end subroutine check_int

   subroutine finish(suite)
      character(*), intent(in) :: suite
      write (*, '(a, a, i5, a, i5, a)') suite, ": ", npass, " passed, ", nfail, " failed"
      if (nfail > 0) error stop 1
   end subroutine finish

end module hrw_check
