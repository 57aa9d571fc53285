program example2_pbe
!! The reference's example 2 on the device (example/example2_pbe_2d_fv.f90:24-69,131-164): the
!! 2D population balance n_t + (g1 n)_x1 + (g2 n)_x2 = 0 with g1 = g2 = 1 on 250 x 250 cells of
!! [0, 10]^2, WENO k = 3 (eps = 1e-6) in both directions, Godunov (upwind) fluxes, zero
!! boundary fluxes, mstvd, dt = 5e-3, the state written at t = 5*i/100, i = 0..100; the initial
!! state is 1 on [1, 3]^2 and 0 elsewhere, u((j-1)*nc1 + i) at cell (i, j).  The right-hand side
!! is the fused device problem.
!!
!! The results go to the tables the reference's plot script reads (example2:165-230): x1.txt
!! and x2.txt (i, x(i), dx(i)) and u.txt (t, u(1..nc1*nc2), one row per output time), same
!! columns and formats.  Usage: example2_pbe [folder [nc1 nc2]]  (default: output/example2,
!! 250 250; the folder must exist).
   use, intrinsic :: iso_fortran_env, only: stdout => output_unit
   use hrweno_kinds, only: rk
   use hrweno_grids, only: grid1
   use hrweno_tvdode, only: mstvd
   use hrweno_problems, only: problem, pbe2d, HRW_BC_ZERO_FLUX
   implicit none

   integer, parameter :: nout = 100
   real(rk), parameter :: tend = 5.0_rk, dt = 5e-3_rk
   integer :: nc(2)
   type(grid1) :: gx(2)
   type(problem) :: prob
   type(mstvd) :: ode
   real(rk), allocatable :: u(:)
   real(rk) :: t
   character(len=:), allocatable :: folder
   integer :: ux, uu, i, j, d, io
   integer(8) :: c0, c1, rate

   folder = arg_or(1, "output/example2")
   nc(1) = int_arg_or(2, 250)
   nc(2) = int_arg_or(3, 250)
   do d = 1, 2
      call gx(d)%linear(xmin=0.0_rk, xmax=10.0_rk, ncells=nc(d))
   end do
   prob = pbe2d(nc, k=3, eps=1e-6_rk, bc=HRW_BC_ZERO_FLUX, x1=[0.0_rk, 10.0_rk], &
                x2=[0.0_rk, 10.0_rk], g1=[1.0_rk, 0.0_rk], g2=[1.0_rk, 0.0_rk])
   ode = mstvd(prob)
   allocate (u(nc(1)*nc(2)))
   do j = 1, nc(2)
      do i = 1, nc(1)
         u((j - 1)*nc(1) + i) = pulse(gx(1)%center(i), gx(2)%center(j))
      end do
   end do

   write (stdout, '(1x, a)') "Running example2 ..."
   call system_clock(c0, rate)
   do d = 1, 2
      open (newunit=ux, file=folder//"/x"//itoa(d)//".txt", status="replace", action="write")
      write (ux, '(a5, 2(1x, a15))') "i", "x(i)", "dx(i)"
      do i = 1, nc(d)
         write (ux, '(i5, 2(1x, es15.5))') i, gx(d)%center(i), gx(d)%width(i)
      end do
      close (ux)
   end do
   open (newunit=uu, file=folder//"/u.txt", status="replace", action="write")
   write (uu, '(a16)', advance="no") "t"
   do i = 1, size(u)
      write (uu, '(1x, a16)', advance="no") "u("//itoa(i)//")"
   end do
   write (uu, '(a)') ""

   t = 0.0_rk
   do io = 0, nout
      call ode%integrate(u, t, tend*io/nout, dt)
      write (uu, '(es16.5e3)', advance="no") t
      do i = 1, size(u)
         write (uu, '(1x, es16.5e3)', advance="no") u(i)
      end do
      write (uu, '(a)') ""
   end do
   close (uu)
   call system_clock(c1)
   write (stdout, '(1x, a, 1x, f5.2)') "Elapsed time (s) :", real(c1 - c0, rk)/real(rate, rk)
   call prob%destroy()

contains

   pure real(rk) function pulse(x1, x2)
   !! 1 on the square [1, 3] x [1, 3], 0 elsewhere.
      real(rk), intent(in) :: x1, x2
      if (x1 >= 1.0_rk .and. x1 <= 3.0_rk .and. x2 >= 1.0_rk .and. x2 <= 3.0_rk) then
         pulse = 1.0_rk
      else
         pulse = 0.0_rk
      end if
   end function pulse

   function itoa(i) result(s)
      integer, intent(in) :: i
      character(len=:), allocatable :: s
      character(len=12) :: b
      write (b, '(i0)') i
      s = trim(b)
   end function itoa

   function arg_or(n, default) result(s)
      integer, intent(in) :: n
      character(*), intent(in) :: default
      character(len=:), allocatable :: s
      character(len=1024) :: b
      integer :: length, stat
      call get_command_argument(n, b, length, stat)
      if (stat == 0 .and. length > 0) then
         s = trim(b)
      else
         s = default
      end if
   end function arg_or

   integer function int_arg_or(n, default)
      integer, intent(in) :: n, default
      character(len=:), allocatable :: s
      integer :: stat
      s = arg_or(n, "")
      int_arg_or = default
      if (len(s) > 0) then
         read (s, *, iostat=stat) int_arg_or
         if (stat /= 0) error stop "the cell counts must be integers"
      end if
   end function int_arg_or

end program example2_pbe
