module hrweno_tvdode
!! TVD Runge-Kutta (orders 1-3) and 5-step third-order multistep integrators with the public
!! API of the reference module hrweno_tvdode (src/hrweno_tvdode.f90): `rktvd(fu, neq, order)`,
!! `mstvd(fu, neq)` and `%integrate(u, t, tout, dt[, itask])`, the fields neq, order, fevals,
!! istate and msg, and the reference's loop control (strict is_done, so the last step may
!! overshoot tout).
!!
!! The time loop and the stage combinations run on the MI355X (libhrweno_hip):
!!  * rktvd(fu, neq, order) / mstvd(fu, neq): `fu` is the user's host routine.  The state
!!    lives on the device; before each call of fu it is copied into a host buffer and udot is
!!    copied back afterwards.
!!  * rktvd(prob, order) / mstvd(prob) with a device problem descriptor (hrweno_problems): the
!!    RHS is fused with each stage update, one device pass per stage.
!!  * rktvd_dev(fu, neq, order) / mstvd_dev(fu, neq): `fu(t, u, udot)` is the user's routine
!!    on device arrays (type(c_ptr)), e.g. composed of weno%reconstruct_dev and the device
!!    fluxes; nothing is copied to the host.
!! `%integrate` takes a host array (copied in and out once per call); `%integrate_dev` takes
!! a device array and copies nothing.
!!
!! An object owns a library ODE, a device buffer and (callback forms) the callback context;
!! copies made by intrinsic assignment share them, counted, and the last copy to go (out of
!! scope, assigned over) releases them.  Integrators built on a problem keep a reference to it.
!! Additions beside the reference API: %save_state / %load_state (the object's state with u and
!! t, for resume in another object or process) and %check_finite (the scan for NaN/Inf that
!! ends every %integrate, on by default: a hit is a fatal error).
   use, intrinsic :: iso_c_binding
   use hrweno_kinds, only: rk
   use hrweno_capi
   use hrweno_problems, only: problem
   implicit none
   private

   public :: rktvd, mstvd, rktvd_dev, mstvd_dev, integrand, integrand_dev

   abstract interface
      subroutine integrand(t, u, udot)
      !! Integrand u'(t, u) (src/hrweno_tvdode.f90:50-57).
         import :: rk
         real(rk), intent(in) :: t, u(:)
         real(rk), intent(out) :: udot(:)
      end subroutine

      subroutine integrand_dev(t, u, udot)
      !! The same on device arrays of neq doubles (work enqueued on the null stream).
         import :: rk, c_ptr
         real(rk), intent(in) :: t
         type(c_ptr), intent(in) :: u, udot
      end subroutine
   end interface

   type :: callback_ctx
      procedure(integrand), pointer, nopass :: fu => null()
      procedure(integrand_dev), pointer, nopass :: fu_dev => null()
      integer :: neq = 0
      real(rk), allocatable :: u(:), udot(:)
   end type callback_ctx

   type :: ode_res
   !! What an integrator owns on the library side, shared by its copies and released by the
   !! last one: the ODE object, the device staging buffer, the callback context and a reference
   !! to the problem it was built on.
      type(c_ptr) :: h = c_null_ptr
      type(c_ptr) :: ubuf = c_null_ptr
      type(callback_ctx), pointer :: ctx => null()
      type(problem) :: prob
      integer, pointer :: refs => null()
   contains
      procedure, private :: ode_res_assign
      generic :: assignment(=) => ode_res_assign
      final :: ode_res_final
   end type ode_res

   type, abstract :: tvdode
   !! Common part of the TVD solvers (src/hrweno_tvdode.f90:14-34).
      procedure(integrand), pointer, nopass, private :: fu => null()
         !! user integrand (callback form)
      integer :: neq
         !! number of equations
      integer :: order
         !! order of the method
      integer :: fevals = 0
         !! number of function evaluations
      integer :: istate = 0
         !! 1 first call for a problem, 2 subsequent call, -1 error
      character(:), allocatable :: msg
         !! error message
      type(ode_res), private :: r
         !! library ODE object, device buffer, callback context (shared, counted)
   contains
      procedure, pass(self) :: error_msg
      procedure, pass(self) :: integrate_dev => tvdode_integrate_dev
      procedure, pass(self) :: check_finite => tvdode_check_finite
      procedure, pass(self) :: save_state => tvdode_save_state
      procedure, pass(self) :: load_state => tvdode_load_state
      procedure, pass(self) :: save_state_dev => tvdode_save_state_dev
      procedure, pass(self) :: load_state_dev => tvdode_load_state_dev
      procedure, pass(self) :: tape => tvdode_tape
      procedure, pass(self) :: adjoint => tvdode_adjoint
      procedure, pass(self) :: adjoint_dev => tvdode_adjoint_dev
   end type tvdode

   type, extends(tvdode) :: rktvd
   !! Runge-Kutta TVD solver.
   contains
      procedure, pass(self) :: integrate => rktvd_integrate
   end type rktvd

   type, extends(tvdode) :: mstvd
   !! Multistep TVD solver.
   contains
      procedure, pass(self) :: integrate => mstvd_integrate
   end type mstvd

   interface rktvd
      module procedure :: rktvd_init, rktvd_init_problem
   end interface rktvd

   interface mstvd
      module procedure :: mstvd_init, mstvd_init_problem
   end interface mstvd

contains

   subroutine ode_res_release(r)
      type(ode_res), intent(inout) :: r
      integer(c_int) :: rc
      if (associated(r%refs)) then
         r%refs = r%refs - 1
         if (r%refs == 0) then
            if (c_associated(r%h)) rc = hrw_ode_destroy(r%h)
            if (c_associated(r%ubuf)) rc = hrw_device_free(r%ubuf)
            if (associated(r%ctx)) deallocate (r%ctx)
            deallocate (r%refs)
         end if
      end if
      nullify (r%refs, r%ctx)
      r%h = c_null_ptr
      r%ubuf = c_null_ptr
      call r%prob%destroy()
   end subroutine ode_res_release

   subroutine ode_res_assign(lhs, rhs)
      class(ode_res), intent(inout) :: lhs
      type(ode_res), intent(in) :: rhs
      if (associated(lhs%refs) .and. associated(rhs%refs)) then
         if (associated(lhs%refs, rhs%refs)) return
      end if
      call ode_res_release(lhs)
      lhs%h = rhs%h
      lhs%ubuf = rhs%ubuf
      lhs%ctx => rhs%ctx
      lhs%prob = rhs%prob
      lhs%refs => rhs%refs
      if (associated(lhs%refs)) lhs%refs = lhs%refs + 1
   end subroutine ode_res_assign

   subroutine ode_res_final(r)
      type(ode_res), intent(inout) :: r
      call ode_res_release(r)
   end subroutine ode_res_final

   integer(c_int) function device_callback(ctx, t, u, udot, stream) bind(C)
   !! Calls the user's device integrand.  The ODE's stream is drained before the call (u is
   !! ready) and the null stream after it (udot is written), so the user's work may use the
   !! null stream.
      type(c_ptr), value :: ctx, u, udot, stream
      real(c_double), value :: t
      type(callback_ctx), pointer :: c
      call c_f_pointer(ctx, c)
      device_callback = hrw_stream_synchronize(stream)
      if (device_callback /= 0) return
      call c%fu_dev(t, u, udot)
      device_callback = hrw_stream_synchronize(c_null_ptr)
   end function device_callback

   integer(c_int) function host_callback(ctx, t, u, udot, stream) bind(C)
   !! Calls the user's host integrand on host copies of the device arrays.
      type(c_ptr), value :: ctx, u, udot, stream
      real(c_double), value :: t
      type(callback_ctx), pointer :: c
      integer(c_size_t) :: bytes
      call c_f_pointer(ctx, c)
      bytes = int(c%neq, c_size_t)*8_c_size_t
      host_callback = hrw_memcpy_dtoh(c_loc(c%u), u, bytes, stream)
      if (host_callback /= 0) return
      call c%fu(t, c%u, c%udot)
      host_callback = hrw_memcpy_htod(udot, c_loc(c%udot), bytes, stream)
   end function host_callback

   subroutine setup_callback(self, fu, neq)
      class(tvdode), intent(inout) :: self
      procedure(integrand) :: fu
      integer, intent(in) :: neq
      self%fu => fu
      allocate (self%r%ctx)
      self%r%ctx%fu => fu
      self%r%ctx%neq = neq
      allocate (self%r%ctx%u(neq), self%r%ctx%udot(neq))
   end subroutine setup_callback

   subroutine setup_callback_dev(self, fu, neq)
      class(tvdode), intent(inout) :: self
      procedure(integrand_dev) :: fu
      integer, intent(in) :: neq
      allocate (self%r%ctx)
      self%r%ctx%fu_dev => fu
      self%r%ctx%neq = neq
   end subroutine setup_callback_dev

   subroutine finish_init(self, rc)
      class(tvdode), intent(inout) :: self
      integer(c_int), intent(in) :: rc
      if (.not. associated(self%r%refs)) then
         allocate (self%r%refs)
         self%r%refs = 1
      end if
      if (rc /= 0) call self%error_msg(hrw_message())
      if (hrw_device_alloc(int(self%neq, c_size_t)*8_c_size_t, self%r%ubuf) /= 0) &
         call self%error_msg(hrw_message())
      self%istate = 1
   end subroutine finish_init

   type(rktvd) function rktvd_init(fu, neq, order) result(self)
   !! rktvd(fu, neq, order) (src/hrweno_tvdode.f90:69-95).
      procedure(integrand) :: fu
      integer, intent(in) :: neq
      integer, intent(in) :: order
      integer(c_int) :: rc
      if (neq > 0) then
         self%neq = neq
      else
         call self%error_msg("Invalid input 'neq'. Valid range: neq >= 1.")
      end if
      if ((order >= 1) .and. (order <= 3)) then
         self%order = order
      else
         call self%error_msg("Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.")
      end if
      call setup_callback(self, fu, neq)
      rc = hrw_rktvd_create_cb(c_funloc(host_callback), c_loc(self%r%ctx), int(neq, c_int64_t), &
                               int(order, c_int), HRW_ARITH_EXACT, self%r%h)
      call finish_init(self, rc)
   end function rktvd_init

   type(rktvd) function rktvd_dev(fu, neq, order) result(self)
   !! rktvd with an integrand on device arrays: the state never leaves the device.
      procedure(integrand_dev) :: fu
      integer, intent(in) :: neq
      integer, intent(in) :: order
      integer(c_int) :: rc
      if (neq > 0) then
         self%neq = neq
      else
         call self%error_msg("Invalid input 'neq'. Valid range: neq >= 1.")
      end if
      if ((order >= 1) .and. (order <= 3)) then
         self%order = order
      else
         call self%error_msg("Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.")
      end if
      call setup_callback_dev(self, fu, neq)
      rc = hrw_rktvd_create_cb(c_funloc(device_callback), c_loc(self%r%ctx), &
                               int(neq, c_int64_t), int(order, c_int), HRW_ARITH_EXACT, self%r%h)
      call finish_init(self, rc)
   end function rktvd_dev

   type(rktvd) function rktvd_init_problem(prob, order) result(self)
   !! rktvd on a device problem descriptor: fused RHS + stage passes.
      type(problem), intent(in) :: prob
      integer, intent(in) :: order
      integer(c_int) :: rc
      self%neq = prob%neq
      if ((order >= 1) .and. (order <= 3)) then
         self%order = order
      else
         call self%error_msg("Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.")
      end if
      self%r%prob = prob
      rc = hrw_rktvd_create(prob%h(), int(order, c_int), self%r%h)
      call finish_init(self, rc)
   end function rktvd_init_problem

   type(mstvd) function mstvd_init(fu, neq) result(self)
   !! mstvd(fu, neq) (src/hrweno_tvdode.f90:180-201).
      procedure(integrand) :: fu
      integer, intent(in) :: neq
      integer(c_int) :: rc
      if (neq > 0) then
         self%neq = neq
      else
         call self%error_msg("Invalid input 'neq'. Valid range: neq >= 1.")
      end if
      self%order = 3
      call setup_callback(self, fu, neq)
      rc = hrw_mstvd_create_cb(c_funloc(host_callback), c_loc(self%r%ctx), int(neq, c_int64_t), &
                               HRW_ARITH_EXACT, self%r%h)
      call finish_init(self, rc)
   end function mstvd_init

   type(mstvd) function mstvd_dev(fu, neq) result(self)
   !! mstvd with an integrand on device arrays.
      procedure(integrand_dev) :: fu
      integer, intent(in) :: neq
      integer(c_int) :: rc
      if (neq > 0) then
         self%neq = neq
      else
         call self%error_msg("Invalid input 'neq'. Valid range: neq >= 1.")
      end if
      self%order = 3
      call setup_callback_dev(self, fu, neq)
      rc = hrw_mstvd_create_cb(c_funloc(device_callback), c_loc(self%r%ctx), &
                               int(neq, c_int64_t), HRW_ARITH_EXACT, self%r%h)
      call finish_init(self, rc)
   end function mstvd_dev

   type(mstvd) function mstvd_init_problem(prob) result(self)
   !! mstvd on a device problem descriptor.
      type(problem), intent(in) :: prob
      integer(c_int) :: rc
      self%neq = prob%neq
      self%order = 3
      self%r%prob = prob
      rc = hrw_mstvd_create(prob%h(), self%r%h)
      call finish_init(self, rc)
   end function mstvd_init_problem

   subroutine tvdode_integrate_dev(self, u, t, tout, dt, itask)
   !! %integrate on a device array u (neq doubles); no host copies.
      class(tvdode), intent(inout) :: self
      type(c_ptr), intent(in) :: u
      real(rk), intent(inout) :: t
      real(rk), intent(in) :: tout, dt
      integer, intent(in), optional :: itask
      integer(c_int) :: task, rc, order, fe, st
      integer(c_int64_t) :: neq
      if (self%istate < 1) return
      task = 1
      if (present(itask)) task = itask
      rc = hrw_ode_integrate(self%r%h, u, t, tout, dt, task)
      if (rc /= 0) call self%error_msg(hrw_message())
      rc = hrw_ode_info(self%r%h, neq, order, fe, st)
      self%fevals = fe
      self%istate = st
   end subroutine tvdode_integrate_dev

   subroutine integrate_host(self, u, t, tout, dt, itask)
      class(tvdode), intent(inout) :: self
      real(rk), intent(inout) :: u(:)
      real(rk), intent(inout) :: t
      real(rk), intent(in) :: tout, dt
      integer, intent(in) :: itask
      real(rk), allocatable, target :: w(:)
      integer(c_size_t) :: bytes
      if (self%istate < 1) return
      if (size(u) /= self%neq) call self%error_msg("Invalid input 'u': size(u) /= neq.")
      w = u
      bytes = int(self%neq, c_size_t)*8_c_size_t
      if (hrw_memcpy_htod(self%r%ubuf, c_loc(w), bytes, c_null_ptr) /= 0) &
         call self%error_msg(hrw_message())
      call self%integrate_dev(self%r%ubuf, t, tout, dt, itask)
      if (hrw_memcpy_dtoh(c_loc(w), self%r%ubuf, bytes, c_null_ptr) /= 0) &
         call self%error_msg(hrw_message())
      u = w
   end subroutine integrate_host

   subroutine rktvd_integrate(self, u, t, tout, dt, itask)
   !! Optimal TVD RK of order 1-3 to tout, or one step if itask = 2
   !! (src/hrweno_tvdode.f90:97-178).
      class(rktvd), intent(inout) :: self
      real(rk), intent(inout) :: u(:)
      real(rk), intent(inout) :: t
      real(rk), intent(in) :: tout
      real(rk), intent(in) :: dt
      integer, intent(in), optional :: itask
      integer :: task
      task = 1
      if (present(itask)) task = itask
      call integrate_host(self, u, t, tout, dt, task)
   end subroutine rktvd_integrate

   subroutine mstvd_integrate(self, u, t, tout, dt)
   !! 5-step third-order TVD multistep to tout, started by four RK3 steps on the first call
   !! (src/hrweno_tvdode.f90:203-271).
      class(mstvd), intent(inout) :: self
      real(rk), intent(inout) :: u(:)
      real(rk), intent(inout) :: t
      real(rk), intent(in) :: tout
      real(rk), intent(in) :: dt
      call integrate_host(self, u, t, tout, dt, 1)
   end subroutine mstvd_integrate

   subroutine tvdode_tape(self, ckpt_every)
   !! Record the following %integrate calls for %adjoint (0 stops recording).  rktvd records
   !! its last %integrate call with a device checkpoint every ckpt_every steps; mstvd keeps
   !! every state from the call that starts the method (istate = 1) on, so it must be set
   !! before that call.  Needs an integrator on a burgers1d problem with a linear grid (either
   !! flux, either kind of ends).  (Added form: the reference has no adjoint.)
      class(tvdode), intent(inout) :: self
      integer, intent(in) :: ckpt_every
      if (hrw_ode_tape(self%r%h, int(ckpt_every, c_int)) /= 0) call self%error_msg(hrw_message())
   end subroutine tvdode_tape

   subroutine tvdode_adjoint_dev(self, lambda)
   !! %adjoint on a device array lambda (neq doubles).
      class(tvdode), intent(inout) :: self
      type(c_ptr), intent(in) :: lambda
      if (hrw_ode_adjoint(self%r%h, lambda) /= 0) call self%error_msg(hrw_message())
   end subroutine tvdode_adjoint_dev

   subroutine tvdode_adjoint(self, lambda)
   !! Discrete adjoint of the recording: lambda holds dJ/du at its final state on entry and
   !! dJ/du at its initial state on return (rktvd: the start of the last recorded %integrate
   !! call; mstvd: the start of the method).
      class(tvdode), intent(inout) :: self
      real(rk), intent(inout) :: lambda(:)
      real(rk), allocatable, target :: w(:)
      integer(c_size_t) :: bytes
      if (size(lambda) /= self%neq) call self%error_msg("Invalid input 'lambda': size(lambda) /= neq.")
      w = lambda
      bytes = int(self%neq, c_size_t)*8_c_size_t
      if (hrw_memcpy_htod(self%r%ubuf, c_loc(w), bytes, c_null_ptr) /= 0) &
         call self%error_msg(hrw_message())
      call self%adjoint_dev(self%r%ubuf)
      if (hrw_memcpy_dtoh(c_loc(w), self%r%ubuf, bytes, c_null_ptr) /= 0) &
         call self%error_msg(hrw_message())
      lambda = w
   end subroutine tvdode_adjoint

   subroutine tvdode_check_finite(self, on)
   !! Switch the NaN/Inf scan that ends every %integrate on (the default) or off.
      class(tvdode), intent(inout) :: self
      logical, intent(in) :: on
      integer(c_int) :: flag
      flag = 0
      if (on) flag = 1
      if (hrw_ode_set_check(self%r%h, flag) /= 0) call self%error_msg(hrw_message())
   end subroutine tvdode_check_finite

   subroutine tvdode_save_state_dev(self, u, t, state)
   !! The object's state (fevals, istate, the multistep history) with the device array u and
   !! t, as bytes (write them to a file to resume in another process).
      class(tvdode), intent(inout) :: self
      type(c_ptr), intent(in) :: u
      real(rk), intent(in) :: t
      integer(c_int8_t), allocatable, target, intent(out) :: state(:)
      integer(c_int64_t) :: nb
      if (hrw_ode_state_size(self%r%h, nb) /= 0) call self%error_msg(hrw_message())
      allocate (state(nb))
      if (hrw_ode_save(self%r%h, u, t, c_loc(state), nb) /= 0) call self%error_msg(hrw_message())
   end subroutine tvdode_save_state_dev

   subroutine tvdode_load_state_dev(self, state, u, t)
   !! Restore a state written by %save_state(_dev) of an integrator of the same kind, order
   !! and neq: the next %integrate continues exactly as the saved object would have.
      class(tvdode), intent(inout) :: self
      integer(c_int8_t), intent(in), target :: state(:)
      type(c_ptr), intent(in) :: u
      real(rk), intent(out) :: t
      integer(c_int64_t) :: neq
      integer(c_int) :: order, fe, st
      if (hrw_ode_load(self%r%h, u, t, c_loc(state), int(size(state), c_int64_t)) /= 0) &
         call self%error_msg(hrw_message())
      if (hrw_ode_info(self%r%h, neq, order, fe, st) /= 0) call self%error_msg(hrw_message())
      self%fevals = fe
      self%istate = st
   end subroutine tvdode_load_state_dev

   subroutine tvdode_save_state(self, u, t, state)
   !! %save_state_dev on a host array u.
      class(tvdode), intent(inout) :: self
      real(rk), intent(in) :: u(:)
      real(rk), intent(in) :: t
      integer(c_int8_t), allocatable, intent(out) :: state(:)
      real(rk), allocatable, target :: w(:)
      if (size(u) /= self%neq) call self%error_msg("Invalid input 'u': size(u) /= neq.")
      w = u
      if (hrw_memcpy_htod(self%r%ubuf, c_loc(w), int(self%neq, c_size_t)*8_c_size_t, &
                          c_null_ptr) /= 0) call self%error_msg(hrw_message())
      call self%save_state_dev(self%r%ubuf, t, state)
   end subroutine tvdode_save_state

   subroutine tvdode_load_state(self, state, u, t)
   !! %load_state_dev into a host array u.
      class(tvdode), intent(inout) :: self
      integer(c_int8_t), intent(in) :: state(:)
      real(rk), intent(out) :: u(:)
      real(rk), intent(out) :: t
      real(rk), allocatable, target :: w(:)
      if (size(u) /= self%neq) call self%error_msg("Invalid input 'u': size(u) /= neq.")
      call self%load_state_dev(state, self%r%ubuf, t)
      allocate (w(self%neq))
      if (hrw_memcpy_dtoh(c_loc(w), self%r%ubuf, int(self%neq, c_size_t)*8_c_size_t, &
                          c_null_ptr) /= 0) call self%error_msg(hrw_message())
      u = w
   end subroutine tvdode_load_state

   pure subroutine error_msg(self, msg)
   !! Error method (src/hrweno_tvdode.f90:286-297).
      class(tvdode), intent(inout) :: self
      character(*), intent(in) :: msg
      self%msg = msg
      self%istate = -1
      error stop self%msg
   end subroutine error_msg

end module hrweno_tvdode
