program test_weno
!! The reference's hrweno suite (test/test_hrweno.f90: same inputs and tolerances) through the
!! Fortran drop-in module, plus strided sections and the device-array form.
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_weno, only: weno, c1, c2, c3, HRW_ARITH_EXACT
   use hrweno_capi, only: hrw_device_alloc, hrw_memcpy_htod, hrw_memcpy_dtoh
   use hrw_check
   implicit none

   call uniform()
   call calc_cnu()
   call nonuniform()
   call strided_and_device()
   call finish("weno")

contains

   subroutine uniform()
      type(weno) :: myweno
      integer, parameter :: nc = 30
      real(rk) :: v(nc), vl(nc), vr(nc)
      integer :: k
      do k = 1, 3
         myweno = weno(nc, k, eps=1e-6_rk)
         v = 0
         v(nc/3:2*nc/3) = 1
         call myweno%reconstruct(v, vl, vr)
         call check_array("uniform vl", vl, v, thr=1e-6_rk)
         call check_array("uniform vr", vr, v, thr=1e-6_rk)
      end do
   end subroutine uniform

   subroutine calc_cnu()
      type(weno) :: myweno
      integer, parameter :: nc = 30
      real(rk) :: xedges(0:nc)
      real(rk), allocatable :: cref(:, :)
      integer :: i, k
      do i = 0, nc
         xedges(i) = 0.0_rk + (3.0_rk - 0.0_rk)*i/nc
      end do
      do k = 1, 3
         myweno = weno(nc, k, eps=1e-6_rk, xedges=xedges)
         select case (k)
         case (1); cref = c1
         case (2); cref = c2
         case (3); cref = c3
         end select
         do i = 1, nc
            call check_array("cnu", reshape(myweno%cnu(:, :, i), [size(cref)]), &
                             reshape(cref, [size(cref)]), thr=1e-5_rk, rel=.true.)
         end do
      end do
   end subroutine calc_cnu

   subroutine nonuniform()
      type(weno) :: myweno
      integer, parameter :: nc = 30
      real(rk) :: v(nc), vl(nc), vr(nc), xedges(0:nc)
      integer :: i, k
      do i = 0, nc
         xedges(i) = 0.0_rk + (1.0_rk - 0.0_rk)*i/nc
      end do
      xedges = xedges**3
      do k = 1, 3
         myweno = weno(nc, k, xedges=xedges)
         v = 0
         v(nc/3:2*nc/3) = 1
         call myweno%reconstruct(v, vl, vr)
         call check_array("nonuniform vl", vl, v, thr=1e-6_rk)
         call check_array("nonuniform vr", vr, v, thr=1e-6_rk)
      end do
   end subroutine nonuniform

   subroutine strided_and_device()
   !! A strided column section (as in example2:107) and the device form give the contiguous
   !! host result bit for bit.
      integer, parameter :: n1 = 7, n2 = 40
      real(rk) :: a(n1*n2), col(n2), l1(n2), r1(n2), l2(n2), r2(n2)
      real(rk), target :: hv(n2), hl(n2), hr(n2)
      type(weno) :: w
      type(c_ptr) :: dv, dl, dr
      integer :: i
      integer(c_size_t) :: bytes
      do i = 1, n1*n2
         a(i) = sin(0.37_rk*i) + merge(1.0_rk, 0.0_rk, mod(i, 11) < 4)
      end do
      w = weno(n2, 3, 1e-6_rk, arith=HRW_ARITH_EXACT)
      col = a(3:3 + (n2 - 1)*n1:n1)
      call w%reconstruct(col, l1, r1)
      call w%reconstruct(a(3:3 + (n2 - 1)*n1:n1), l2, r2)
      call check_array("strided vl", l2, l1, thr=0.0_rk)
      call check_array("strided vr", r2, r1, thr=0.0_rk)
      bytes = int(n2, c_size_t)*8
      if (hrw_device_alloc(bytes, dv) /= 0) error stop "alloc"
      if (hrw_device_alloc(bytes, dl) /= 0) error stop "alloc"
      if (hrw_device_alloc(bytes, dr) /= 0) error stop "alloc"
      hv = col
      if (hrw_memcpy_htod(dv, c_loc(hv), bytes, c_null_ptr) /= 0) error stop "htod"
      call w%reconstruct_dev(dv, dl, dr)
      if (hrw_memcpy_dtoh(c_loc(hl), dl, bytes, c_null_ptr) /= 0) error stop "dtoh"
      if (hrw_memcpy_dtoh(c_loc(hr), dr, bytes, c_null_ptr) /= 0) error stop "dtoh"
      call check_array("device vl", hl, l1, thr=0.0_rk)
      call check_array("device vr", hr, r1, thr=0.0_rk)
   end subroutine strided_and_device

end program test_weno
