program example1_burgers
!! The reference's example 1 on the device (example/example1_burgers_1d_fv.f90:29-66,120-133):
!! Burgers' equation u_t + (u**2/2)_x = 0 on 100 cells of [-5, 5], WENO k = 3 (eps = 1e-6),
!! Godunov interface fluxes with zero-gradient boundary fluxes, rktvd order 3, dt = 1e-2, the
!! state written at t = 12*i/100, i = 0..100.  The right-hand side is the fused device problem
!! (one device pass per stage) instead of a user callback.
!!
!! The results go to the tables the reference's plot script reads (example1:135-188): x.txt
!! (i, x(i), dx(i)) and u.txt (t, u(1..nc), one row per output time), same columns and
!! formats.  Usage: example1_burgers [folder]  (default: output/example1; it must exist).
   use, intrinsic :: iso_fortran_env, only: stdout => output_unit
   use hrweno_kinds, only: rk
   use hrweno_grids, only: grid1
   use hrweno_tvdode, only: rktvd
   use hrweno_problems, only: problem, burgers1d, HRW_GODUNOV, HRW_BC_ZERO_GRADIENT
   implicit none

   integer, parameter :: nc = 100, nout = 100
   real(rk), parameter :: tend = 12.0_rk, dt = 1e-2_rk
   type(grid1) :: gx
   type(problem) :: prob
   type(rktvd) :: ode
   real(rk) :: u(nc), t
   character(len=:), allocatable :: folder
   integer :: ux, uu, i, io
   integer(8) :: c0, c1, rate

   folder = arg_or(1, "output/example1")
   call gx%linear(xmin=-5.0_rk, xmax=5.0_rk, ncells=nc)
   prob = burgers1d(nc, k=3, eps=1e-6_rk, flux=HRW_GODUNOV, bc=HRW_BC_ZERO_GRADIENT, &
                    xmin=-5.0_rk, xmax=5.0_rk)
   ode = rktvd(prob, 3)
   do i = 1, nc
      u(i) = ramp(gx%center(i))
   end do

   write (stdout, '(1x, a)') "Running example1 ..."
   call system_clock(c0, rate)
   open (newunit=ux, file=folder//"/x.txt", status="replace", action="write")
   write (ux, '(a5, 2(1x, a15))') "i", "x(i)", "dx(i)"
   do i = 1, nc
      write (ux, '(i5, 2(1x, es15.5))') i, gx%center(i), gx%width(i)
   end do
   close (ux)
   open (newunit=uu, file=folder//"/u.txt", status="replace", action="write")
   write (uu, '(a16)', advance="no") "t"
   do i = 1, nc
      write (uu, '(1x, a16)', advance="no") "u("//itoa(i)//")"
   end do
   write (uu, '(a)') ""

   t = 0.0_rk
   do io = 0, nout
      call ode%integrate(u, t, tend*io/nout, dt)
      write (uu, '(es16.5e3)', advance="no") t
      do i = 1, nc
         write (uu, '(1x, es16.5e3)', advance="no") u(i)
      end do
      write (uu, '(a)') ""
   end do
   close (uu)
   call system_clock(c1)
   write (stdout, '(1x, a, 1x, f6.1)') "Elapsed time (ms) :", 1e3_rk*real(c1 - c0, rk)/real(rate, rk)
   call prob%destroy()

contains

   elemental real(rk) function ramp(x)
   !! A linear profile from 1 at x = -4 to -0.5 at x = 2, held constant outside.
      real(rk), intent(in) :: x
      real(rk), parameter :: xa = -4.0_rk, xb = 2.0_rk, ua = 1.0_rk, ub = -0.5_rk
      ramp = max(min(ua + (ub - ua)/(xb - xa)*(x - xa), ua), ub)
   end function ramp

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

end program example1_burgers
