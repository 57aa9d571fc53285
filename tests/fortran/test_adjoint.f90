program test_adjoint
!! The added %adjoint method of rktvd on a fused periodic Lax-Friedrichs Burgers problem:
!!  * the gradient of J = sum(u(tout)**2)/2 with respect to u(t0) agrees with central finite
!!    differences (h = 1e-6) of the same %integrate calls at several cells (rel 1e-6), on a
!!    problem with a fixed Lax-Friedrichs alpha (with alpha = max|u| the cell holding the
!!    maximum sits on a kink of J, where a finite difference averages two one-sided slopes);
!!  * for J = sum(width*u(tout)), conserved by periodic finite volumes, the gradient is width;
!!  * the checkpoint interval does not change the result.
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_grids, only: grid1
   use hrweno_tvdode, only: rktvd
   use hrweno_problems, only: problem, burgers1d
   use hrw_check
   implicit none

   integer, parameter :: nc = 200
   real(rk), parameter :: tout = 0.15_rk, h = 1e-6_rk
   type(grid1) :: gx
   type(problem) :: prob, prob_fixed
   type(rktvd) :: ode, ode_fixed
   real(rk) :: u0(nc), u(nc), lam(nc), lam2(nc), up(nc), um(nc), t, dt, fd
   integer :: i, j, cells(6)

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

end program test_adjoint
