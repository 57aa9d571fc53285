program test_adjoint
!! The added %adjoint method on fused Burgers problems (rktvd on periodic Lax-Friedrichs lines;
!! mstvd on example 1's own composition, Godunov with zero-gradient ends):
!!  * the gradient of J = sum(u(tout)**2)/2 with respect to u(t0) agrees with central finite
!!    differences (h = 1e-6) of the same %integrate calls at several cells (rel 1e-6), on a
!!    problem with a fixed Lax-Friedrichs alpha (with alpha = max|u| the cell holding the
!!    maximum sits on a kink of J, where a finite difference averages two one-sided slopes);
!!  * for J = sum(width*u(tout)), conserved by periodic finite volumes, the gradient is width;
!!  * the checkpoint interval does not change the result;
!!  * mstvd, recorded from the start of the method over two %integrate calls: the gradient of
!!    J = sum(u(tout)**2)/2 against central differences at several cells, and the end cells,
!!    which zero-gradient ends never change, get exactly their own cotangent back.
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_grids, only: grid1
   use hrweno_tvdode, only: rktvd, mstvd
   use hrweno_problems, only: problem, burgers1d, HRW_GODUNOV, HRW_BC_ZERO_GRADIENT
   use hrw_check
   implicit none

   integer, parameter :: nc = 200
   real(rk), parameter :: tout = 0.15_rk, h = 1e-6_rk
   type(grid1) :: gx
   type(problem) :: prob, prob_fixed, prob_ex1
   type(rktvd) :: ode, ode_fixed
   type(mstvd) :: ms
   real(rk) :: u0(nc), u(nc), lam(nc), lam2(nc), up(nc), um(nc), t, dt, fd
   integer :: i, j, cells(6)
   integer, parameter :: ms_cells(5) = [17, 60, 101, 120, 180]

   call gx%linear(xmin=-1.0_rk, xmax=1.0_rk, ncells=nc)
   do i = 1, nc
      u0(i) = 0.5_rk + sin(3.141592653589793_rk*gx%center(i))
      if (gx%center(i) > 0.1_rk) u0(i) = u0(i) + 0.4_rk
   end do
   dt = 0.4_rk*(2.0_rk/nc)/maxval(abs(u0))
   prob = burgers1d(nc, k=3)
   prob_fixed = burgers1d(nc, k=3, alpha=2.5_rk)

   ! gradient of 0.5*sum(u(tout)**2)
   ode_fixed = rktvd(prob_fixed, 3)
   call ode_fixed%tape(7)
   u = u0
   t = 0.0_rk
   call ode_fixed%integrate(u, t, tout, dt)
   lam = u
   call ode_fixed%adjoint(lam)

   cells = [1, 17, 60, 101, 150, 200]
   do i = 1, size(cells)
      j = cells(i)
      up = u0
      up(j) = up(j) + h
      um = u0
      um(j) = um(j) - h
      call advance(up)
      call advance(um)
      fd = sum((up - um)*(up + um))/(4.0_rk*h)
      call check("dJ/du0 vs FD", lam(j), fd, thr=1e-6_rk, rel=.true.)
   end do

   ! alpha = max|u|: the same with every step checkpointed is identical
   ode = rktvd(prob, 3)
   call ode%tape(7)
   u = u0
   t = 0.0_rk
   call ode%integrate(u, t, tout, dt)
   lam = u
   call ode%adjoint(lam)
   call ode%tape(1)
   u = u0
   t = 0.0_rk
   call ode%integrate(u, t, tout, dt)
   lam2 = u
   call ode%adjoint(lam2)
   call check_array("checkpoint interval", lam2, lam, thr=0.0_rk)

   ! conserved total: the gradient of sum(width*u) is width
   u = u0
   t = 0.0_rk
   call ode%integrate(u, t, tout, dt)
   lam = gx%width
   call ode%adjoint(lam)
   call check_array("conserved total", lam, gx%width, thr=1e-12_rk*maxval(gx%width))

   ! mstvd on example 1's composition, recorded over two calls from the start of the method
   prob_ex1 = burgers1d(nc, k=3, flux=HRW_GODUNOV, bc=HRW_BC_ZERO_GRADIENT)
   ms = mstvd(prob_ex1)
   call ms%tape(1)
   u = u0
   t = 0.0_rk
   call ms%integrate(u, t, tout/2, dt)
   call ms%integrate(u, t, tout, dt)
   lam = u
   call ms%adjoint(lam)
   ! (cells away from the shock that forms at x = 0.1: around it the endpoint Godunov flux has
   ! kinks where |vm| = |vp| (src/hrweno_fluxes.f90:66-74) within 1e-6 of u0, which a
   ! difference of that size straddles -- cell 150 is one: its difference settles on the
   ! adjoint's value only at h = 1e-7)
   do i = 1, size(ms_cells)
      j = ms_cells(i)
      ! Richardson-extrapolated central differences, (4 D(h) - D(2h))/3
      fd = (4.0_rk*ms_diff(j, h) - ms_diff(j, 2.0_rk*h))/3.0_rk
      call check("mstvd dJ/du0 vs FD", lam(j), fd, thr=1e-6_rk, rel=.true.)
   end do
   lam = 0.0_rk
   lam(nc) = 1.0_rk
   call ms%adjoint(lam)
   call check("mstvd fixed end cell", lam(nc), 1.0_rk, thr=4e-16_rk)
   call check("mstvd fixed end cell, elsewhere", maxval(abs(lam(1:nc - 1))), 0.0_rk, thr=0.0_rk)
   call finish("adjoint")

contains

   subroutine advance(v)
      real(rk), intent(inout) :: v(:)
      type(rktvd) :: o
      real(rk) :: tt
      o = rktvd(prob_fixed, 3)
      tt = 0.0_rk
      call o%integrate(v, tt, tout, dt)
   end subroutine advance

   real(rk) function ms_diff(j, step)
   !! central difference of J = sum(u(tout)**2)/2 in u0(j) through the mstvd run
      integer, intent(in) :: j
      real(rk), intent(in) :: step
      real(rk) :: vp(nc), vm(nc)
      vp = u0
      vp(j) = vp(j) + step
      vm = u0
      vm(j) = vm(j) - step
      call advance_ms(vp)
      call advance_ms(vm)
      ms_diff = sum((vp - vm)*(vp + vm))/(4.0_rk*step)
   end function ms_diff

   subroutine advance_ms(v)
      real(rk), intent(inout) :: v(:)
      type(mstvd) :: o
      real(rk) :: tt
      o = mstvd(prob_ex1)
      tt = 0.0_rk
      call o%integrate(v, tt, tout/2, dt)
      call o%integrate(v, tt, tout, dt)
   end subroutine advance_ms

end program test_adjoint
