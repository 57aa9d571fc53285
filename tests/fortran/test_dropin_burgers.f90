program test_dropin_burgers
!! Example 1 through the drop-in API (example/example1_burgers_1d_fv.f90:72-109: a user RHS
!! calling weno%reconstruct and godunov, zero-gradient boundary fluxes, rktvd order 3) equals
!! the fused device problem (burgers1d, HRW_GODUNOV, HRW_BC_ZERO_GRADIENT) bit for bit in the
!! EXACT arithmetic form; the host-array and device-array integrate forms agree too.
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_weno, only: weno, HRW_ARITH_EXACT
   use hrweno_fluxes, only: godunov
   use hrweno_grids, only: grid1
   use hrweno_tvdode, only: rktvd, mstvd
   use hrweno_problems, only: problem, burgers1d, HRW_GODUNOV, HRW_BC_ZERO_GRADIENT
   use hrweno_capi, only: hrw_device_alloc, hrw_memcpy_htod, hrw_memcpy_dtoh
   use hrw_check
   implicit none

   integer, parameter :: nc = 100
   type(grid1) :: gx
   type(weno) :: myweno
   type(rktvd) :: ode_cb, ode_fused, ode_dev
   type(mstvd) :: ms_cb, ms_fused
   type(problem) :: prob
   real(rk) :: u_cb(nc), u_fused(nc), u0(nc)
   real(rk), target :: hbuf(nc)
   real(rk) :: t_cb, t_fused, t_dev, dt
   type(c_ptr) :: du
   integer :: ii, i

   call gx%linear(xmin=-5.0_rk, xmax=5.0_rk, ncells=nc)
   myweno = weno(ncells=nc, k=3, eps=1e-6_rk, arith=HRW_ARITH_EXACT)
   do i = 1, nc
      u0(i) = ic(gx%center(i))
   end do
   dt = 1e-2_rk

   ode_cb = rktvd(rhs, nc, order=3)
   prob = burgers1d(nc, k=3, eps=1e-6_rk, flux=HRW_GODUNOV, bc=HRW_BC_ZERO_GRADIENT, &
                    xmin=-5.0_rk, xmax=5.0_rk, arith=HRW_ARITH_EXACT)
   ode_fused = rktvd(prob, 3)
   ode_dev = rktvd(prob, 3)
   u_cb = u0
   u_fused = u0
   t_cb = 0.0_rk
   t_fused = 0.0_rk
   t_dev = 0.0_rk
   if (hrw_device_alloc(int(nc, c_size_t)*8, du) /= 0) error stop "alloc"
   hbuf = u0
   if (hrw_memcpy_htod(du, c_loc(hbuf), int(nc, c_size_t)*8, c_null_ptr) /= 0) error stop "htod"
   do ii = 1, 20
      call ode_cb%integrate(u_cb, t_cb, 0.3_rk*ii, dt)
      call ode_fused%integrate(u_fused, t_fused, 0.3_rk*ii, dt)
      call ode_dev%integrate_dev(du, t_dev, 0.3_rk*ii, dt)
   end do
   if (hrw_memcpy_dtoh(c_loc(hbuf), du, int(nc, c_size_t)*8, c_null_ptr) /= 0) error stop "dtoh"
   call check("t", t_fused, t_cb, thr=0.0_rk)
   call check("t dev", t_dev, t_cb, thr=0.0_rk)
   call check_int("fevals", ode_fused%fevals, ode_cb%fevals)
   call check_array("callback == fused", u_fused, u_cb, thr=0.0_rk)
   call check_array("device == host", hbuf, u_fused, thr=0.0_rk)

   ! multistep: callback path vs fused path
   ms_cb = mstvd(rhs, nc)
   ms_fused = mstvd(prob)
   u_cb = u0
   u_fused = u0
   t_cb = 0.0_rk
   t_fused = 0.0_rk
   call ms_cb%integrate(u_cb, t_cb, 1.0_rk, dt)
   call ms_fused%integrate(u_fused, t_fused, 1.0_rk, dt)
   call check("ms t", t_fused, t_cb, thr=0.0_rk)
   call check_int("ms fevals", ms_fused%fevals, ms_cb%fevals)
   call check_array("ms callback == fused", u_fused, u_cb, thr=0.0_rk)

   ! non-uniform grid: weno(..., xedges) in the callback equals burgers1d(..., xedges=)
   call gx%geometric(xmin=1.0_rk, xmax=11.0_rk, ratio=1.02_rk, ncells=nc)
   myweno = weno(ncells=nc, k=3, eps=1e-6_rk, xedges=gx%edges, arith=HRW_ARITH_EXACT)
   call prob%destroy()
   prob = burgers1d(nc, k=3, eps=1e-6_rk, flux=HRW_GODUNOV, bc=HRW_BC_ZERO_GRADIENT, &
                    xedges=gx%edges, arith=HRW_ARITH_EXACT)
   ode_cb = rktvd(rhs, nc, order=3)
   ode_fused = rktvd(prob, 3)
   do i = 1, nc
      u_cb(i) = ic(gx%center(i) - 6.0_rk)
   end do
   u_fused = u_cb
   t_cb = 0.0_rk
   t_fused = 0.0_rk
   call ode_cb%integrate(u_cb, t_cb, 2.0_rk, dt)
   call ode_fused%integrate(u_fused, t_fused, 2.0_rk, dt)
   call check("nonuniform t", t_fused, t_cb, thr=0.0_rk)
   call check_array("nonuniform callback == fused", u_fused, u_cb, thr=0.0_rk)
   call finish("dropin_burgers")

contains

   subroutine rhs(t, v, vdot)
   !! example/example1_burgers_1d_fv.f90:72-109
      real(rk), intent(in) :: t
      real(rk), intent(in) :: v(:)
      real(rk), intent(out) :: vdot(:)
      real(rk) :: fedges(0:nc), vl(nc), vr(nc)
      integer :: i
      call myweno%reconstruct(v, vl, vr)
      do i = 1, nc - 1
         fedges(i) = godunov(flux, vr(i), vl(i + 1), [gx%right(i)], t)
      end do
      fedges(0) = fedges(1)
      fedges(nc) = fedges(nc - 1)
      vdot = -(fedges(1:) - fedges(:nc - 1))/gx%width
   end subroutine rhs

   pure real(rk) function flux(v, x, t)
      real(rk), intent(in) :: v, x(:), t
      flux = (v**2)/2
   end function flux

   elemental real(rk) function ic(x)
      real(rk), intent(in) :: x
      real(rk), parameter :: xa = -4.0_rk, xb = 2.0_rk, va = 1.0_rk, vb = -0.5_rk
      ic = va + (vb - va)/(xb - xa)*(x - xa)
      ic = max(min(ic, va), vb)
   end function ic

end program test_dropin_burgers
