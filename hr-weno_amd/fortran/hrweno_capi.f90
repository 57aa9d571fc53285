module hrweno_capi
!! iso_c_binding interfaces to libhrweno_hip.so (include/hrweno_capi.h).  Functions that
!! the drop-in modules call from pure procedures are declared pure: they touch only their
!! arguments and device memory owned by the library.
   use, intrinsic :: iso_c_binding
   implicit none
   public

   integer(c_int), parameter :: HRW_OK = 0
   integer(c_int), parameter :: HRW_ARITH_EXACT = 0, HRW_ARITH_FAST = 1
   integer(c_int), parameter :: HRW_LAX_FRIEDRICHS = 0, HRW_GODUNOV = 1   ! hrw_numflux
   integer(c_int), parameter :: HRW_FN_BURGERS = 0, HRW_FN_LINEAR = 1
   integer(c_int), parameter :: HRW_BC_PERIODIC = 0, HRW_BC_ZERO_GRADIENT = 1, &
                                HRW_BC_ZERO_FLUX = 2, HRW_BC_INFLOW0_OUTFLOW = 3

   interface
      pure integer(c_int) function hrw_last_error_copy(buf, len) bind(C)
         import :: c_int, c_char, c_int64_t
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int64_t), value :: len
      end function

      ! ---- WENO
      pure integer(c_int) function hrw_weno_calc_cnu(ncells, k, xedges, cnu) bind(C)
         import :: c_int, c_double
         integer(c_int), value :: ncells, k
         real(c_double), intent(in) :: xedges(*)
         real(c_double), intent(out) :: cnu(*)
      end function

      pure integer(c_int) function hrw_reconstruct_host(ncells, k, eps, arith, cnu, v, incv, &
                                                        vl, incl, vr, incr) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: ncells, k, arith
         real(c_double), value :: eps
         type(c_ptr), value :: cnu
         real(c_double), intent(in) :: v(*)
         integer(c_int64_t), value :: incv, incl, incr
         real(c_double), intent(out) :: vl(*), vr(*)
      end function

      integer(c_int) function hrw_reconstruct_dev(ncells, k, eps, arith, v, incv, vl, incl, &
                                                  vr, incr, stream) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: ncells, k, arith
         real(c_double), value :: eps
         type(c_ptr), value :: v, vl, vr, stream
         integer(c_int64_t), value :: incv, incl, incr
      end function

      integer(c_int) function hrw_reconstruct_lines_host(ncells, k, eps, arith, cnu, nlines, &
                                                         v, incv, ldv, vl, vr, inco, ldo) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: ncells, k, arith
         real(c_double), value :: eps
         type(c_ptr), value :: cnu
         integer(c_int64_t), value :: nlines, incv, ldv, inco, ldo
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(inout) :: vl(*), vr(*)
      end function

      integer(c_int) function hrw_reconstruct_lines_dev(ncells, k, eps, arith, nlines, v, incv, &
                                                        ldv, vl, vr, inco, ldo, stream) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: ncells, k, arith
         real(c_double), value :: eps
         integer(c_int64_t), value :: nlines, incv, ldv, inco, ldo
         type(c_ptr), value :: v, vl, vr, stream
      end function

      ! ---- fluxes on device arrays
      integer(c_int) function hrw_flux_lax_friedrichs(fn, a, b, n, vm, vp, x, t, alpha, h, &
                                                      stream) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: fn
         real(c_double), value :: a, b, t, alpha
         integer(c_int64_t), value :: n
         type(c_ptr), value :: vm, vp, x, h, stream
      end function

      integer(c_int) function hrw_flux_godunov(fn, a, b, n, vm, vp, x, t, h, stream) bind(C)
         import :: c_int, c_double, c_ptr, c_int64_t
         integer(c_int), value :: fn
         real(c_double), value :: a, b, t
         integer(c_int64_t), value :: n
         type(c_ptr), value :: vm, vp, x, h, stream
      end function

      ! ---- grids (host arrays)
      pure integer(c_int) function hrw_grid_linear(xmin, xmax, n, edges) bind(C)
         import :: c_int, c_double
         real(c_double), value :: xmin, xmax
         integer(c_int), value :: n
         real(c_double), intent(out) :: edges(*)
      end function

      pure integer(c_int) function hrw_grid_bilinear(xmin, xcross, xmax, n1, n2, edges) bind(C)
         import :: c_int, c_double
         real(c_double), value :: xmin, xcross, xmax
         integer(c_int), value :: n1, n2
         real(c_double), intent(out) :: edges(*)
      end function

      pure integer(c_int) function hrw_grid_log(xmin, xmax, n, edges) bind(C)
         import :: c_int, c_double
         real(c_double), value :: xmin, xmax
         integer(c_int), value :: n
         real(c_double), intent(out) :: edges(*)
      end function

      integer(c_int) function hrw_grid_geometric(xmin, xmax, ratio, n, edges) bind(C)
         import :: c_int, c_double
         real(c_double), value :: xmin, xmax, ratio
         integer(c_int), value :: n
         real(c_double), intent(out) :: edges(*)
      end function

      pure integer(c_int) function hrw_grid_compute(n, edges, center, width) bind(C)
         import :: c_int, c_double
         integer(c_int), value :: n
         real(c_double), intent(in) :: edges(*)
         real(c_double), intent(out) :: center(*), width(*)
      end function

      ! ---- device memory and streams
      integer(c_int) function hrw_device_alloc(bytes, ptr) bind(C)
         import :: c_int, c_size_t, c_ptr
         integer(c_size_t), value :: bytes
         type(c_ptr), intent(out) :: ptr
      end function

      integer(c_int) function hrw_device_free(ptr) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: ptr
      end function

      integer(c_int) function hrw_device_mem_info(free_bytes, total_bytes) bind(C)
         import :: c_int, c_size_t
         integer(c_size_t), intent(out) :: free_bytes, total_bytes
      end function

      integer(c_int) function hrw_memcpy_htod(dst, src, bytes, stream) bind(C)
         import :: c_int, c_size_t, c_ptr
         type(c_ptr), value :: dst, src, stream
         integer(c_size_t), value :: bytes
      end function

      integer(c_int) function hrw_memcpy_dtoh(dst, src, bytes, stream) bind(C)
         import :: c_int, c_size_t, c_ptr
         type(c_ptr), value :: dst, src, stream
         integer(c_size_t), value :: bytes
      end function

      integer(c_int) function hrw_stream_synchronize(stream) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: stream
      end function

      ! ---- problems and integrators
      integer(c_int) function hrw_problem_burgers1d(n, k, eps, numflux, bc, alpha, xmin, xmax, &
                                                    width, arith, out) bind(C)
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: n, k, numflux, bc, arith
         real(c_double), value :: eps, alpha, xmin, xmax
         type(c_ptr), value :: width
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_problem_burgers1d_edges(n, k, eps, numflux, bc, alpha, &
                                                          xedges, arith, out) bind(C)
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: n, k, numflux, bc, arith
         real(c_double), value :: eps, alpha
         type(c_ptr), value :: xedges
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_problem_pbe2d(n1, n2, k, eps, bc, x1min, x1max, x2min, &
                                                x2max, g1a, g1b, g2a, g2b, arith, out) bind(C)
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: n1, n2, k, bc, arith
         real(c_double), value :: eps, x1min, x1max, x2min, x2max, g1a, g1b, g2a, g2b
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_problem_destroy(p) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: p
      end function

      integer(c_int) function hrw_problem_rhs(p, t, u, udot, stream) bind(C)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: p, u, udot, stream
         real(c_double), value :: t
      end function

      integer(c_int) function hrw_rktvd_create(p, order, out) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: p
         integer(c_int), value :: order
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_rktvd_create_cb(fu, ctx, neq, order, arith, out) bind(C)
         import :: c_int, c_ptr, c_funptr, c_int64_t
         type(c_funptr), value :: fu
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: neq
         integer(c_int), value :: order, arith
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_mstvd_create(p, out) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: p
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_mstvd_create_cb(fu, ctx, neq, arith, out) bind(C)
         import :: c_int, c_ptr, c_funptr, c_int64_t
         type(c_funptr), value :: fu
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: neq
         integer(c_int), value :: arith
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function hrw_ode_destroy(ode) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: ode
      end function

      integer(c_int) function hrw_ode_integrate(ode, u, t, tout, dt, itask) bind(C)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ode, u
         real(c_double), intent(inout) :: t
         real(c_double), value :: tout, dt
         integer(c_int), value :: itask
      end function

      integer(c_int) function hrw_ode_info(ode, neq, order, fevals, istate) bind(C)
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ode
         integer(c_int64_t), intent(out) :: neq
         integer(c_int), intent(out) :: order, fevals, istate
      end function

      integer(c_int) function hrw_ode_set_check(ode, check_finite) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: ode
         integer(c_int), value :: check_finite
      end function

      integer(c_int) function hrw_ode_state_size(ode, bytes) bind(C)
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ode
         integer(c_int64_t), intent(out) :: bytes
      end function

      integer(c_int) function hrw_ode_save(ode, u, t, buf, bytes) bind(C)
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ode, u, buf
         real(c_double), value :: t
         integer(c_int64_t), value :: bytes
      end function

      integer(c_int) function hrw_ode_load(ode, u, t, buf, bytes) bind(C)
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: ode, u, buf
         real(c_double), intent(out) :: t
         integer(c_int64_t), value :: bytes
      end function

      integer(c_int) function hrw_ode_tape(ode, ckpt_every) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: ode
         integer(c_int), value :: ckpt_every
      end function

      integer(c_int) function hrw_ode_adjoint(ode, lambda) bind(C)
         import :: c_int, c_ptr
         type(c_ptr), value :: ode, lambda
      end function
   end interface

contains

   pure function hrw_message() result(msg)
   !! Message of the last failed library call on this thread.
      character(:), allocatable :: msg
      character(kind=c_char) :: buf(512)
      integer :: i, n, rc
      rc = hrw_last_error_copy(buf, 512_c_int64_t)
      n = 0
      do i = 1, size(buf)
         if (buf(i) == c_null_char) exit
         n = i
      end do
      allocate (character(n) :: msg)
      do i = 1, n
         msg(i:i) = buf(i)
      end do
   end function hrw_message

end module hrweno_capi
