module hrweno_problems
!! Device problem descriptors: the user right-hand sides of the reference examples
!! (example/example1_burgers_1d_fv.f90:72-122, example/example2_pbe_2d_fv.f90:73-155) as
!! objects the integrators of hrweno_tvdode can fuse with their stage updates, one device
!! pass per stage.
   use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_ptr, c_null_ptr, c_loc, &
                                          c_associated
   use hrweno_kinds, only: rk
   use hrweno_capi
   implicit none
   private

   public :: problem, burgers1d, pbe2d
   public :: HRW_LAX_FRIEDRICHS, HRW_GODUNOV, HRW_BC_PERIODIC, HRW_BC_ZERO_GRADIENT, &
             HRW_BC_ZERO_FLUX, HRW_BC_INFLOW0_OUTFLOW, HRW_ARITH_EXACT, HRW_ARITH_FAST

   type :: problem_ref
   !! Shared ownership of one library problem: copies (intrinsic assignment of a problem, or of
   !! an integrator built from it) count references, and the last one to go destroys it.
      type(c_ptr) :: h = c_null_ptr
      integer, pointer :: refs => null()
   contains
      procedure, private :: problem_ref_assign
      generic :: assignment(=) => problem_ref_assign
      final :: problem_ref_final
   end type problem_ref

   type :: problem
      type(problem_ref), private :: ref
         !! library handle (shared, counted)
      integer :: neq = 0
         !! number of equations (cells)
   contains
      procedure, pass(self) :: h => problem_handle
      procedure, pass(self) :: rhs_dev => problem_rhs_dev
      procedure, pass(self) :: destroy => problem_destroy
   end type problem

contains

   subroutine problem_ref_release(r)
      type(problem_ref), intent(inout) :: r
      integer(c_int) :: rc
      if (associated(r%refs)) then
         r%refs = r%refs - 1
         if (r%refs == 0) then
            if (c_associated(r%h)) rc = hrw_problem_destroy(r%h)
            deallocate (r%refs)
         end if
      end if
      nullify (r%refs)
      r%h = c_null_ptr
   end subroutine problem_ref_release

   subroutine problem_ref_assign(lhs, rhs)
      class(problem_ref), intent(inout) :: lhs
      type(problem_ref), intent(in) :: rhs
      if (associated(lhs%refs) .and. associated(rhs%refs)) then
         if (associated(lhs%refs, rhs%refs)) return
      end if
      call problem_ref_release(lhs)
      lhs%h = rhs%h
      lhs%refs => rhs%refs
      if (associated(lhs%refs)) lhs%refs = lhs%refs + 1
   end subroutine problem_ref_assign

   subroutine problem_ref_final(r)
      type(problem_ref), intent(inout) :: r
      call problem_ref_release(r)
   end subroutine problem_ref_final

   subroutine adopt(p, h)
   !! p takes the first reference to a newly created library problem.
      type(problem), intent(inout) :: p
      type(c_ptr), intent(in) :: h
      p%ref%h = h
      allocate (p%ref%refs)
      p%ref%refs = 1
   end subroutine adopt

   type(c_ptr) function problem_handle(self) result(h)
   !! The library handle (valid while this object or an integrator built from it lives).
      class(problem), intent(in) :: self
      h = self%ref%h
   end function problem_handle

   function burgers1d(ncells, k, eps, flux, bc, alpha, xmin, xmax, width, arith, xedges) result(p)
   !! 1D Burgers u_t + (u**2/2)_x = 0 on a linear grid [xmin, xmax] (or given cell widths).
   !! flux: HRW_LAX_FRIEDRICHS (alpha < 0: max|v| of each stage input) or
   !! HRW_GODUNOV; bc: HRW_BC_PERIODIC or HRW_BC_ZERO_GRADIENT (example 1).
   !! xedges(0:ncells): a non-uniform grid whose own WENO coefficients the reconstruction uses,
   !! as weno(ncells, k, eps, xedges) (src/hrweno_weno.f90:100-112); needs HRW_BC_ZERO_GRADIENT.
      integer, intent(in) :: ncells
      integer, intent(in), optional :: k, flux, bc, arith
      real(rk), intent(in), optional :: eps, alpha, xmin, xmax
      real(rk), intent(in), optional, target :: width(:), xedges(:)
      type(problem) :: p
      integer :: k_, flux_, bc_, arith_
      real(rk) :: eps_, alpha_, xmin_, xmax_
      type(c_ptr) :: pw, h
      k_ = 3; flux_ = HRW_LAX_FRIEDRICHS; bc_ = HRW_BC_PERIODIC; arith_ = HRW_ARITH_FAST
      eps_ = 1e-6_rk; alpha_ = -1.0_rk; xmin_ = -1.0_rk; xmax_ = 1.0_rk
      if (present(k)) k_ = k
      if (present(flux)) flux_ = flux
      if (present(bc)) bc_ = bc
      if (present(arith)) arith_ = arith
      if (present(eps)) eps_ = eps
      if (present(alpha)) alpha_ = alpha
      if (present(xmin)) xmin_ = xmin
      if (present(xmax)) xmax_ = xmax
      if (present(xedges)) then
         if (size(xedges) /= ncells + 1) error stop "Invalid input 'xedges': size(xedges) /= ncells + 1."
         if (.not. is_contiguous(xedges)) error stop "Invalid input 'xedges': not contiguous."
         if (hrw_problem_burgers1d_edges(ncells, k_, eps_, flux_, bc_, alpha_, c_loc(xedges), &
                                         arith_, h) /= 0) error stop hrw_message()
         call adopt(p, h)
         p%neq = ncells
         return
      end if
      pw = c_null_ptr
      if (present(width)) then
         if (size(width) /= ncells) error stop "Invalid input 'width': size(width) /= ncells."
         if (.not. is_contiguous(width)) error stop "Invalid input 'width': not contiguous."
         pw = c_loc(width)
      end if
      if (hrw_problem_burgers1d(ncells, k_, eps_, flux_, bc_, alpha_, xmin_, xmax_, pw, arith_, &
                                h) /= 0) error stop hrw_message()
      call adopt(p, h)
      p%neq = ncells
   end function burgers1d

   function pbe2d(ncells, k, eps, bc, x1, x2, g1, g2, arith) result(p)
   !! 2D population balance n_t + (g1 n)_x1 + (g2 n)_x2 = 0 on linear grids, growth rates
   !! g1 = g1(1) + g1(2)*x1 and g2 = g2(1) + g2(2)*x2, Godunov flux, u((j-1)*ncells(1)+i).
   !! bc: HRW_BC_ZERO_FLUX (example 2) or HRW_BC_INFLOW0_OUTFLOW.
      integer, intent(in) :: ncells(2)
      integer, intent(in), optional :: k, bc, arith
      real(rk), intent(in), optional :: eps, x1(2), x2(2), g1(2), g2(2)
      type(problem) :: p
      integer :: k_, bc_, arith_
      real(rk) :: eps_, x1_(2), x2_(2), g1_(2), g2_(2)
      type(c_ptr) :: h
      k_ = 3; bc_ = HRW_BC_INFLOW0_OUTFLOW; arith_ = HRW_ARITH_FAST; eps_ = 1e-6_rk
      x1_ = [0.0_rk, 1.0_rk]; x2_ = [0.0_rk, 1.0_rk]
      g1_ = [1.0_rk, 0.5_rk]; g2_ = [0.5_rk, 0.0_rk]
      if (present(k)) k_ = k
      if (present(bc)) bc_ = bc
      if (present(arith)) arith_ = arith
      if (present(eps)) eps_ = eps
      if (present(x1)) x1_ = x1
      if (present(x2)) x2_ = x2
      if (present(g1)) g1_ = g1
      if (present(g2)) g2_ = g2
      if (hrw_problem_pbe2d(ncells(1), ncells(2), k_, eps_, bc_, x1_(1), x1_(2), x2_(1), x2_(2), &
                            g1_(1), g1_(2), g2_(1), g2_(2), arith_, h) /= 0) &
         error stop hrw_message()
      call adopt(p, h)
      p%neq = ncells(1)*ncells(2)
   end function pbe2d

   subroutine problem_rhs_dev(self, t, u, udot, stream)
   !! udot = L(u) on device arrays (one right-hand-side evaluation).
      class(problem), intent(in) :: self
      real(rk), intent(in) :: t
      type(c_ptr), intent(in) :: u, udot
      type(c_ptr), intent(in), optional :: stream
      type(c_ptr) :: s
      s = c_null_ptr
      if (present(stream)) s = stream
      if (hrw_problem_rhs(self%ref%h, t, u, udot, s) /= 0) error stop hrw_message()
   end subroutine problem_rhs_dev

   subroutine problem_destroy(self)
   !! Drop this object's reference to the device descriptor (it is destroyed with the last
   !! reference: integrators built from it keep their own).  Going out of scope or being
   !! assigned over does the same.
      class(problem), intent(inout) :: self
      call problem_ref_release(self%ref)
      self%neq = 0
   end subroutine problem_destroy

end module hrweno_problems
