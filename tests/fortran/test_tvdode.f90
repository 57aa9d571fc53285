program test_tvdode
!! The reference's tvdode suite (test/test_tvdode.f90: u' = a*u, same inputs and tolerances)
!! through the Fortran drop-in module, plus the closed forms u_N = u0*R(a*dt)**N.
   use hrweno_kinds, only: rk
   use hrweno_tvdode, only: rktvd, mstvd
   use hrw_check
   implicit none
   integer, parameter :: nu = 10
   integer :: ii
   real(rk), parameter :: a(nu) = [(-1.0_rk + real(ii - 1, rk)*4/(nu - 1), ii=1, nu)]

   call rk_orders()
   call ms()
   call finish("tvdode")

contains

   subroutine fu(t, u, udot)
      real(rk), intent(in) :: t, u(:)
      real(rk), intent(out) :: udot(:)
      udot = a*u
   end subroutine

   subroutine rk_orders()
      type(rktvd) :: ode
      integer :: order, nsteps
      real(rk) :: t0, t, tout, dt, u(nu), u0(nu), z(nu), r(nu), rtol(3)
      t0 = -1.3_rk
      tout = 1.5_rk
      u0 = 0.1_rk
      dt = (tout - t0)/3000
      rtol = [2e-2_rk, 1e-3_rk, 1e-3_rk]
      do order = 1, 3
         ode = rktvd(fu, nu, order)
         call check_int("istate before", ode%istate, 1)
         u = u0
         t = t0
         call ode%integrate(u, t, tout, dt)
         call check_array("rk vs exp", u, u0*exp(a*(t - t0)), thr=rtol(order), rel=.true.)
         nsteps = ode%fevals/order
         z = a*dt
         select case (order)
         case (1); r = 1 + z
         case (2); r = 1 + z + z**2/2
         case (3); r = 1 + z + z**2/2 + z**3/6
         end select
         call check_array("rk closed form", u, u0*r**nsteps, thr=1e-10_rk, rel=.true.)
         call check_int("istate after", ode%istate, 2)
         ! single step (itask = 2)
         call ode%integrate(u, t, tout + 1, dt, itask=2)
         call check_int("itask=2 fevals", ode%fevals, (nsteps + 1)*order)
      end do
   end subroutine rk_orders

   subroutine ms()
      type(mstvd) :: ode
      real(rk) :: t0, t, tout, dt, u(nu), u0(nu)
      t0 = -1.3_rk
      tout = 1.5_rk
      u0 = 0.1_rk
      dt = 1e-3_rk
      ode = mstvd(fu, nu)
      u = u0
      t = t0
      call ode%integrate(u, t, tout, dt)
      call check_array("ms vs exp", u, u0*exp(a*(t - t0)), thr=1e-3_rk, rel=.true.)
      call check("ms overshoot", t - tout, 0.5_rk*dt, thr=0.5_rk*dt*1.0001_rk)
      ! startup steps ignore tout: four RK3 steps even when tout is reached earlier
      ode = mstvd(fu, nu)
      u = u0
      t = 0.0_rk
      call ode%integrate(u, t, 0.01_rk, 0.1_rk)
      call check_int("ms startup fevals", ode%fevals, 12)
      call check("ms startup t", t, 0.4_rk, thr=1e-15_rk)
   end subroutine ms

end program test_tvdode
