! ref_harness.f90 -- bind(C) entry points over the reference's OWN compiled modules
! (hrweno_kinds, hrweno_weno, hrweno_fluxes, compiled from the reference checkout by
! oracle/Makefile into oracle/_ref/).  Test infrastructure only: it lets the pytest
! suite compare the C restatement in hrweno_oracle.c against the reference Fortran
! arithmetic bit for bit, and lets bench.py time the reference reconstruction on the
! host.  Nothing in the shipped library loads it.
!
! The reference's hrweno_tvdode and hrweno_grids need fortran-lang/stdlib (optval),
! which is not present in this image, so they are not compiled here.
module hrweno_ref_harness
   use, intrinsic :: iso_c_binding, only: c_int, c_double, c_int64_t
   use hrweno_kinds, only: rk
   use hrweno_weno, only: weno
   use hrweno_fluxes, only: lax_friedrichs, godunov
   implicit none

   real(rk), save :: g_coef = 1.0_rk

contains

   !> weno(n,k,eps[,xedges]) followed by %reconstruct(v,vl,vr)
   !> (src/hrweno_weno.f90:54-127, 129-219).  has_edges=0: uniform grid.
   subroutine ref_weno_reconstruct(n, k, eps, has_edges, xedges, v, vl, vr) &
      bind(C, name="ref_weno_reconstruct")
      integer(c_int), value :: n, k, has_edges
      real(c_double), value :: eps
      real(c_double), intent(in) :: xedges(0:n), v(n)
      real(c_double), intent(out) :: vl(n), vr(n)
      type(weno) :: w
      if (has_edges /= 0) then
         w = weno(n, k, eps, xedges)
      else
         w = weno(n, k, eps)
      end if
      call w%reconstruct(v, vl, vr)
   end subroutine

   !> the public cnu table of weno(n,k,eps,xedges) (src/hrweno_weno.f90:221-297)
   subroutine ref_weno_cnu(n, k, xedges, cnu) bind(C, name="ref_weno_cnu")
      integer(c_int), value :: n, k
      real(c_double), intent(in) :: xedges(0:n)
      real(c_double), intent(out) :: cnu(0:k - 1, -1:k - 1, 1:n)
      type(weno) :: w
      w = weno(n, k, 1e-6_rk, xedges)
      cnu = w%cnu
   end subroutine

   !> repeated reconstruction of one line, for host timing of the reference
   subroutine ref_weno_reconstruct_repeat(n, k, eps, nrep, v, vl, vr) &
      bind(C, name="ref_weno_reconstruct_repeat")
      integer(c_int), value :: n, k, nrep
      real(c_double), value :: eps
      real(c_double), intent(in) :: v(n)
      real(c_double), intent(out) :: vl(n), vr(n)
      type(weno) :: w
      integer :: r
      w = weno(n, k, eps)
      do r = 1, nrep
         call w%reconstruct(v, vl, vr)
      end do
   end subroutine

   pure real(rk) function flux_burgers(u, x, t)
      real(rk), intent(in) :: u, x(:), t
      flux_burgers = (u**2)/2
   end function

   pure real(rk) function flux_linear(u, x, t)
      real(rk), intent(in) :: u, x(:), t
      flux_linear = u*x(1)*t
   end function

   !> lax_friedrichs / godunov over arrays of interface states
   !> (src/hrweno_fluxes.f90:22-76).  kind 0: Burgers u**2/2; kind 1: u*x*t.
   subroutine ref_fluxes(kind, which, n, vm, vp, x, t, alpha, h) bind(C, name="ref_fluxes")
      integer(c_int), value :: kind, which, n
      real(c_double), value :: t, alpha
      real(c_double), intent(in) :: vm(n), vp(n), x(n)
      real(c_double), intent(out) :: h(n)
      integer :: i
      do i = 1, n
         if (kind == 0) then
            if (which == 0) then
               h(i) = lax_friedrichs(flux_burgers, vm(i), vp(i), [x(i)], t, alpha)
            else
               h(i) = godunov(flux_burgers, vm(i), vp(i), [x(i)], t)
            end if
         else
            if (which == 0) then
               h(i) = lax_friedrichs(flux_linear, vm(i), vp(i), [x(i)], t, alpha)
            else
               h(i) = godunov(flux_linear, vm(i), vp(i), [x(i)], t)
            end if
         end if
      end do
   end subroutine

   !> Periodic 1D Burgers, WENO(k) + Lax-Friedrichs (alpha = maxval(abs(v)) per RHS call),
   !> nsteps SSP-RK3 steps of size dt, u advanced in place.  The RHS is the composition of
   !> example/example1_burgers_1d_fv.f90:72-109 with k periodic ghost cells per side, calling
   !> the reference's own weno%reconstruct and lax_friedrichs; the stage lines restate
   !> src/hrweno_tvdode.f90:162-167 (that module needs stdlib and is not compiled here).
   subroutine ref_burgers_rk3(n, k, eps, nsteps, dt, width, u) bind(C, name="ref_burgers_rk3")
      integer(c_int), value :: n, k, nsteps
      real(c_double), value :: eps, dt
      real(c_double), intent(in) :: width(n)
      real(c_double), intent(inout) :: u(n)
      real(rk), allocatable :: ui(:), udot(:), vpad(:), vl(:), vr(:), fedges(:)
      type(weno) :: w
      integer :: s
      w = weno(n + 2*k, k, eps)
      allocate (ui(n), udot(n), vpad(n + 2*k), vl(n + 2*k), vr(n + 2*k), fedges(0:n))
      do s = 1, nsteps
         call rhs(u, udot)
         ui = u + dt*udot
         call rhs(ui, udot)
         ui = (3*u + ui + dt*udot)/4
         call rhs(ui, udot)
         u = (u + 2*ui + 2*dt*udot)/3
      end do
   contains
      subroutine rhs(v, vdot)
         real(rk), intent(in) :: v(:)
         real(rk), intent(out) :: vdot(:)
         real(rk) :: alpha
         integer :: i
         vpad(1:k) = v(n - k + 1:n)
         vpad(k + 1:k + n) = v
         vpad(k + n + 1:) = v(1:k)
         call w%reconstruct(vpad, vl, vr)
         alpha = maxval(abs(v))
         do i = 0, n
            fedges(i) = lax_friedrichs(flux_burgers, vr(k + i), vl(k + i + 1), [0.0_rk], 0.0_rk, alpha)
         end do
         vdot = -(fedges(1:) - fedges(:n - 1))/width
      end subroutine
   end subroutine

   !> 1D Burgers with the composition of example/example1_burgers_1d_fv.f90:72-109 on the
   !> reference's own weno%reconstruct and fluxes: bc 1 reconstructs the line itself (constant
   !> extrapolation inside reconstruct), takes the flux at the n-1 interior interfaces and copies
   !> the end ones (:97-104); bc 0 pads k periodic ghost cells per side.  flux 0:
   !> lax_friedrichs with alpha = maxval(abs(v)), 1: godunov.  method 3: nsteps SSP-RK3 steps
   !> (the stage lines of src/hrweno_tvdode.f90:162-167); method 0: the multistep method from
   !> a fresh object, four RK3 starting steps then nsteps-4 steps of src/hrweno_tvdode.f90:
   !> 236-266 (restated, as that module needs stdlib).  u advanced in place.
   subroutine ref_burgers_ex(n, k, eps, flux, bc, method, nsteps, dt, width, u) &
      bind(C, name="ref_burgers_ex")
      integer(c_int), value :: n, k, flux, bc, method, nsteps
      real(c_double), value :: eps, dt
      real(c_double), intent(in) :: width(n)
      real(c_double), intent(inout) :: u(n)
      real(rk), allocatable :: ui(:), udot(:), vpad(:), vl(:), vr(:), fedges(:), uold(:, :), &
                               udotold(:, :)
      type(weno) :: w
      real(rk) :: alpha
      integer :: s, c
      if (bc == 0) then
         w = weno(n + 2*k, k, eps)
         allocate (vpad(n + 2*k), vl(n + 2*k), vr(n + 2*k))
      else
         w = weno(n, k, eps)
         allocate (vl(n), vr(n))
      end if
      allocate (ui(n), udot(n), fedges(0:n), uold(n, 4), udotold(n, 4))
      if (method == 3) then
         do s = 1, nsteps
            call rk3_step()
         end do
      else
         do c = 4, 1, -1
            uold(:, c) = u
            call rhs(u, udotold(:, c))
            call rk3_step()
         end do
         do s = 5, nsteps
            call rhs(u, udot)
            ui = (25*u + 50*dt*udot + 7*uold(:, 4) + 10*dt*udotold(:, 4))/32
            udotold = eoshift(udotold, shift=-1, dim=2)
            uold = eoshift(uold, shift=-1, dim=2)
            udotold(:, 1) = udot
            uold(:, 1) = u
            u = ui
         end do
      end if
   contains
      subroutine rk3_step()
         call rhs(u, udot)
         ui = u + dt*udot
         call rhs(ui, udot)
         ui = (3*u + ui + dt*udot)/4
         call rhs(ui, udot)
         u = (u + 2*ui + 2*dt*udot)/3
      end subroutine
      subroutine rhs(v, vdot)
         real(rk), intent(in) :: v(:)
         real(rk), intent(out) :: vdot(:)
         integer :: i
         alpha = maxval(abs(v))
         if (bc == 0) then
            vpad(1:k) = v(n - k + 1:n)
            vpad(k + 1:k + n) = v
            vpad(k + n + 1:) = v(1:k)
            call w%reconstruct(vpad, vl, vr)
            do i = 0, n
               fedges(i) = iface(vr(k + i), vl(k + i + 1))
            end do
         else
            call w%reconstruct(v, vl, vr)
            do i = 1, n - 1
               fedges(i) = iface(vr(i), vl(i + 1))
            end do
            fedges(0) = fedges(1)
            fedges(n) = fedges(n - 1)
         end if
         vdot = -(fedges(1:) - fedges(:n - 1))/width
      end subroutine
      real(rk) function iface(vm, vp)
         real(rk), intent(in) :: vm, vp
         if (flux == 0) then
            iface = lax_friedrichs(flux_burgers, vm, vp, [0.0_rk], 0.0_rk, alpha)
         else
            iface = godunov(flux_burgers, vm, vp, [0.0_rk], 0.0_rk)
         end if
      end function
   end subroutine

end module hrweno_ref_harness
