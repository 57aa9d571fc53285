module hrweno_fluxes
!! Monotone numerical fluxes with the public API of the reference module hrweno_fluxes
!! (src/hrweno_fluxes.f90).  `lax_friedrichs` and `godunov` keep the reference's scalar form
!! with a user flux procedure: they are the building blocks of a user's host RHS.  For the
!! built-in flux functions (HRW_FN_BURGERS: u**2/2, HRW_FN_LINEAR: (a + b*x)*u) the
!! `*_dev` forms evaluate whole arrays of interfaces on the device; the fused integrators
!! evaluate them inside their stage kernels.
   use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_ptr, c_null_ptr
   use hrweno_kinds, only: rk
   use hrweno_capi, only: HRW_FN_BURGERS, HRW_FN_LINEAR, hrw_flux_lax_friedrichs, &
                          hrw_flux_godunov, hrw_message
   implicit none
   private

   public :: lax_friedrichs, godunov, lax_friedrichs_dev, godunov_dev, flux
   public :: HRW_FN_BURGERS, HRW_FN_LINEAR

   abstract interface
      pure function flux(u, x, t)
      !! Flux function f(u, x, t) (src/hrweno_fluxes.f90:12-18).
         import :: rk
         real(rk) :: flux
         real(rk), intent(in) :: u, x(:), t
      end function
   end interface

contains

   pure real(rk) function lax_friedrichs(f, vm, vp, x, t, alpha) result(res)
   !! Lax-Friedrichs flux (f(vm) + f(vp) - alpha*(vp - vm))/2, Shu eq. 2.72
   !! (src/hrweno_fluxes.f90:22-45).
      procedure(flux) :: f
      real(rk), intent(in) :: vm, vp, x(:), t, alpha
      real(rk) :: fm, fp
      fm = f(vm, x, t)
      fp = f(vp, x, t)
      res = (fm + fp - alpha*(vp - vm))/2
   end function lax_friedrichs

   pure real(rk) function godunov(f, vm, vp, x, t) result(res)
   !! Godunov-type flux from the end-point values, Shu eq. 2.70 (src/hrweno_fluxes.f90:47-76):
   !! min(f(vm), f(vp)) if vm <= vp, else max(f(vm), f(vp)).
      procedure(flux) :: f
      real(rk), intent(in) :: vm, vp, x(:), t
      real(rk) :: fm, fp
      fm = f(vm, x, t)
      fp = f(vp, x, t)
      if (vm <= vp) then
         res = min(fm, fp)
      else
         res = max(fm, fp)
      end if
   end function godunov

   subroutine lax_friedrichs_dev(fn, n, vm, vp, h, alpha, x, t, coef, stream)
   !! h(i) = lax_friedrichs(fn, vm(i), vp(i), [x(i)], t, alpha) on device arrays.
      integer, intent(in) :: fn, n
      type(c_ptr), intent(in) :: vm, vp, h
      real(rk), intent(in) :: alpha
      type(c_ptr), intent(in), optional :: x, stream
      real(rk), intent(in), optional :: t, coef(2)
      real(rk) :: c(2), tt
      type(c_ptr) :: px, s
      integer(c_int) :: rc
      c = [0.0_rk, 1.0_rk]
      if (present(coef)) c = coef
      tt = 0.0_rk
      if (present(t)) tt = t
      px = c_null_ptr
      if (present(x)) px = x
      s = c_null_ptr
      if (present(stream)) s = stream
      rc = hrw_flux_lax_friedrichs(fn, c(1), c(2), int(n, c_int64_t), vm, vp, px, tt, alpha, h, s)
      if (rc /= 0) error stop hrw_message()
   end subroutine lax_friedrichs_dev

   subroutine godunov_dev(fn, n, vm, vp, h, x, t, coef, stream)
   !! h(i) = godunov(fn, vm(i), vp(i), [x(i)], t) on device arrays.
      integer, intent(in) :: fn, n
      type(c_ptr), intent(in) :: vm, vp, h
      type(c_ptr), intent(in), optional :: x, stream
      real(rk), intent(in), optional :: t, coef(2)
      real(rk) :: c(2), tt
      type(c_ptr) :: px, s
      integer(c_int) :: rc
      c = [0.0_rk, 1.0_rk]
      if (present(coef)) c = coef
      tt = 0.0_rk
      if (present(t)) tt = t
      px = c_null_ptr
      if (present(x)) px = x
      s = c_null_ptr
      if (present(stream)) s = stream
      rc = hrw_flux_godunov(fn, c(1), c(2), int(n, c_int64_t), vm, vp, px, tt, h, s)
      if (rc /= 0) error stop hrw_message()
   end subroutine godunov_dev

end module hrweno_fluxes
