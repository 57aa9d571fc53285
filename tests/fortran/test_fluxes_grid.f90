program test_fluxes_grid
!! The reference's fluxes and grid suites (test/test_fluxes.f90, test/test_grid.f90: same inputs
!! and tolerances) through the Fortran drop-in modules.  Host-only: runs without a GPU.
   use hrweno_kinds, only: rk
   use hrweno_fluxes, only: godunov, lax_friedrichs
   use hrweno_grids, only: grid1
   use hrw_check
   implicit none
   real(rk), parameter :: rtol = 1e-5_rk

   call fluxes()
   call linear()
   call log_grid()
   call geometric()
   call bilinear()
   call finish("fluxes+grid")

contains

   pure function f(u, x, t)
      real(rk) :: f
      real(rk), intent(in) :: u, x(:), t
      f = u*x(1)*t
   end function

   subroutine fluxes()
      real(rk) :: h, href, vm, vp, x(1), t
      vm = 2.0_rk
      x = 3.0_rk
      t = 5.0_rk
      vp = vm
      href = f(vp, x, t)
      h = godunov(f, vm, vp, x, t)
      call check("h(a,a) godunov", h, href, thr=1e-8_rk, rel=.true.)
      h = lax_friedrichs(f, vm, vp, x, t, x(1)*t)
      call check("h(a,a) lf", h, href, thr=1e-8_rk, rel=.true.)
      vp = -2*vm
      href = f(vm, x, t)
      call check("-> godunov", godunov(f, vm, vp, x, t), href, thr=1e-8_rk, rel=.true.)
      call check("-> lf", lax_friedrichs(f, vm, vp, x, t, x(1)*t), href, thr=1e-8_rk, rel=.true.)
      vm = -vm
      vp = -vp
      href = f(vm, x, t)
      call check("<- godunov", godunov(f, vm, vp, x, t), href, thr=1e-8_rk, rel=.true.)
      call check("<- lf", lax_friedrichs(f, vm, vp, x, t, x(1)*t), href, thr=1e-8_rk, rel=.true.)
   end subroutine fluxes

   subroutine common(gx, nc, xmin, xmax)
      type(grid1), intent(in) :: gx
      integer, intent(in) :: nc
      real(rk), intent(in) :: xmin, xmax
      call check_int("ncells", gx%ncells, nc)
      if (xmin /= 0) call check("edges(0)", gx%edges(0), xmin, thr=rtol, rel=.true.)
      call check("edges(nc)", gx%edges(nc), xmax, thr=rtol, rel=.true.)
      call check_array("left", gx%left, gx%edges(0:nc - 1), thr=0.0_rk)
      call check_array("right", gx%right, gx%edges(1:nc), thr=0.0_rk)
      call check_array("width", gx%width, gx%right - gx%left, thr=0.0_rk)
      call check_array("center", gx%center, (gx%left + gx%right)/2, thr=0.0_rk)
   end subroutine common

   subroutine linear()
      type(grid1) :: gx
      call gx%linear(1.0_rk, 1e3_rk, 10**6, name="F [N]")
      call common(gx, 10**6, 1.0_rk, 1e3_rk)
   end subroutine linear

   subroutine log_grid()
      type(grid1) :: gx
      call gx%log(1e-1_rk, 1e3_rk, 10**4, name="T [K]")
      call common(gx, 10**4, 1e-1_rk, 1e3_rk)
   end subroutine log_grid

   subroutine geometric()
      type(grid1) :: gx
      call gx%geometric(1e1_rk, 1e3_rk, 1.1_rk, 10**2, name="P [W]")
      call common(gx, 10**2, 1e1_rk, 1e3_rk)
   end subroutine geometric

   subroutine bilinear()
      type(grid1) :: gx
      call gx%bilinear(0.0_rk, 1e1_rk, 1e3_rk, [124, 365], name="W [J]")
      call common(gx, 489, 0.0_rk, 1e3_rk)
      call check("edges(124)", gx%edges(124), 1e1_rk, thr=rtol, rel=.true.)
      if (gx%name /= "W [J]" .or. gx%scale /= "bilinear") call check_int("name/scale", 0, 1)
   end subroutine bilinear

end program test_fluxes_grid
