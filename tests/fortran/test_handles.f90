program test_handles
!! Ownership and the additions of the drop-in integrators, on the device:
!!  * 1000 integrators created and dropped (assigned over, or going out of scope), in every
!!    form, leave device memory where it was; copies share their library objects safely;
!!    an integrator built on a temporary problem keeps the problem alive;
!!  * rktvd_dev / mstvd_dev (a device-array integrand) equal the fused problem bit for bit;
!!  * %save_state -> fresh object -> %load_state resumes a multistep run bit for bit;
!!  * %reconstruct_lines over example 2's x1 and x2 lines equals %reconstruct line by line.
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_weno, only: weno, HRW_ARITH_EXACT
   use hrweno_tvdode, only: rktvd, mstvd, rktvd_dev, mstvd_dev
   use hrweno_problems, only: problem, burgers1d
   use hrweno_capi, only: hrw_device_mem_info, hrw_device_alloc, hrw_device_free, &
                          hrw_memcpy_htod, hrw_memcpy_dtoh
   use hrw_check
   implicit none

   integer, parameter :: nc = 2000
   type(problem) :: prob
   type(rktvd) :: a, b
   type(mstvd) :: m1, m2, m3
   real(rk) :: u0(nc), u1(nc), u2(nc), t1, t2, dt
   integer(c_size_t) :: free0, free1, total
   integer(c_int8_t), allocatable :: state(:)
   integer :: i

   do i = 1, nc
      u0(i) = 0.5_rk + sin(3.141592653589793_rk*(-1.0_rk + (i - 0.5_rk)*2.0_rk/nc))
   end do
   dt = 0.4_rk*(2.0_rk/nc)/1.5_rk
   prob = burgers1d(nc, arith=HRW_ARITH_EXACT)

   ! ---- ownership: warm up once (allocator pools), then 1000 create/drop cycles
   a = rktvd(prob, 3)
   b = rktvd(rhs, nc, 3)
   m1 = mstvd(prob)
   if (hrw_device_mem_info(free0, total) /= 0) error stop "mem info"
   do i = 1, 1000
      a = rktvd(prob, 3)
      b = rktvd(rhs, nc, 3)
      m1 = mstvd(burgers1d(nc, arith=HRW_ARITH_EXACT))   ! problem temporary kept by m1
      call scoped()
   end do
   if (hrw_device_mem_info(free1, total) /= 0) error stop "mem info"
   call check("device memory flat after 1000 objects (MiB)", &
              real(free0, rk)/2.0_rk**20, real(free1, rk)/2.0_rk**20, thr=1.0_rk)
   ! copies share one library object; the last one alive still works
   b = a
   call a%integrate(u1, t1, 0.0_rk, dt)
   u1 = u0; t1 = 0.0_rk
   call b%integrate(u1, t1, 5.5_rk*dt, dt)
   call check_int("copy integrates", b%istate, 2)
   ! m1 was built on a temporary problem: it still integrates like the named one
   m2 = mstvd(prob)
   u1 = u0; u2 = u0; t1 = 0.0_rk; t2 = 0.0_rk
   call m1%integrate(u1, t1, 12.5_rk*dt, dt)
   call m2%integrate(u2, t2, 12.5_rk*dt, dt)
   call check_array("temporary problem kept alive", u1, u2, thr=0.0_rk)

   ! ---- device-array integrand == fused problem
   a = rktvd(prob, 3)
   b = rktvd_dev(rhs_dev, nc, 3)
   u1 = u0; u2 = u0; t1 = 0.0_rk; t2 = 0.0_rk
   call a%integrate(u1, t1, 20.5_rk*dt, dt)
   call b%integrate(u2, t2, 20.5_rk*dt, dt)
   call check("rktvd_dev t", t2, t1, thr=0.0_rk)
   call check_int("rktvd_dev fevals", b%fevals, a%fevals)
   call check_array("rktvd_dev == fused", u2, u1, thr=0.0_rk)
   m1 = mstvd(prob)
   m2 = mstvd_dev(rhs_dev, nc)
   u1 = u0; u2 = u0; t1 = 0.0_rk; t2 = 0.0_rk
   call m1%integrate(u1, t1, 20.5_rk*dt, dt)
   call m2%integrate(u2, t2, 20.5_rk*dt, dt)
   call check_array("mstvd_dev == fused", u2, u1, thr=0.0_rk)

   ! ---- save, fresh object, load: the split run equals the unbroken one
   m1 = mstvd(prob)
   u1 = u0; t1 = 0.0_rk
   call m1%integrate(u1, t1, 9.5_rk*dt, dt)
   call m1%integrate(u1, t1, 17.5_rk*dt, dt)
   m2 = mstvd(prob)
   u2 = u0; t2 = 0.0_rk
   call m2%integrate(u2, t2, 9.5_rk*dt, dt)
   call m2%save_state(u2, t2, state)
   m3 = mstvd(prob)
   u2 = 0.0_rk
   call m3%load_state(state, u2, t2)
   call m3%integrate(u2, t2, 17.5_rk*dt, dt)
   call check("resumed t", t2, t1, thr=0.0_rk)
   call check_int("resumed fevals", m3%fevals, m1%fevals)
   call check_array("resumed == unbroken", u2, u1, thr=0.0_rk)

   call lines_check()
   call finish("handles")

contains

   subroutine scoped()
      type(rktvd) :: c
      type(mstvd) :: d
      c = rktvd(rhs, nc, 2)
      d = mstvd(rhs, nc)
   end subroutine scoped

   subroutine rhs(t, v, vdot)
      real(rk), intent(in) :: t, v(:)
      real(rk), intent(out) :: vdot(:)
      vdot = -v
   end subroutine rhs

   subroutine rhs_dev(t, v, vdot)
   !! A device-resident integrand: the problem's own right-hand side on device arrays.
      real(rk), intent(in) :: t
      type(c_ptr), intent(in) :: v, vdot
      call prob%rhs_dev(t, v, vdot)
   end subroutine rhs_dev

   subroutine lines_check()
      integer, parameter :: n1 = 37, n2 = 23
      type(weno) :: w1, w2
      real(rk) :: v(n1*n2), vl(n1*n2), vr(n1*n2), al(n1*n2), ar(n1*n2)
      real(rk) :: l1(n1), r1(n1), l2(n2), r2(n2)
      integer :: i, j
      do j = 1, n2
         do i = 1, n1
            v((j - 1)*n1 + i) = exp(-((i - 15.0_rk)**2 + (j - 9.0_rk)**2)/30.0_rk) + &
                                merge(0.3_rk, 0.0_rk, i > 25)
         end do
      end do
      w1 = weno(ncells=n1, k=3, eps=1e-6_rk, arith=HRW_ARITH_EXACT)
      w2 = weno(ncells=n2, k=3, eps=1e-6_rk, arith=HRW_ARITH_EXACT)
      ! x1 lines (example2:97-98): contiguous, one per j
      call w1%reconstruct_lines(v, vl, vr, nlines=n2, incv=1, ldv=n1, inco=1, ldo=n1)
      do j = 1, n2
         call w1%reconstruct(v(1 + (j - 1)*n1:j*n1), l1, r1)
         al(1 + (j - 1)*n1:j*n1) = l1
         ar(1 + (j - 1)*n1:j*n1) = r1
      end do
      call check_array("x1 lines vl", vl, al, thr=0.0_rk)
      call check_array("x1 lines vr", vr, ar, thr=0.0_rk)
      ! x2 lines (example2:106-107): stride n1, one per i, results in the same layout
      call w2%reconstruct_lines(v, vl, vr, nlines=n1, incv=n1, ldv=1, inco=n1, ldo=1)
      do i = 1, n1
         call w2%reconstruct(v(i:i + (n2 - 1)*n1:n1), l2, r2)
         al(i:i + (n2 - 1)*n1:n1) = l2
         ar(i:i + (n2 - 1)*n1:n1) = r2
      end do
      call check_array("x2 lines vl", vl, al, thr=0.0_rk)
      call check_array("x2 lines vr", vr, ar, thr=0.0_rk)
   end subroutine lines_check

end program test_handles
