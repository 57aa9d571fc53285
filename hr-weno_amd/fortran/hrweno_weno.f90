module hrweno_weno
!! WENO reconstruction of order 2k-1 (k = 1, 2, 3) on uniform or non-uniform grids, with the
!! public API of the reference module hrweno_weno (src/hrweno_weno.f90): the `weno`
!! constructor, `%reconstruct(v, vl, vr)`, and the coefficient tables c1, c2, c3.  The
!! reconstruction runs on the MI355X through libhrweno_hip (include/hrweno_capi.h); the
!! non-uniform coefficients cnu are computed there too and kept here as a public array, as in
!! the reference.
!!
!! Additions, all optional: `arith` in the constructor selects the kernels' arithmetic form
!! (HRW_ARITH_FAST, the default, or HRW_ARITH_EXACT, bit-identical to the reference) and
!! `%reconstruct_dev` works on device arrays.
   use, intrinsic :: iso_c_binding, only: c_int, c_int64_t, c_ptr, c_null_ptr, c_loc
   use hrweno_kinds, only: rk
   use hrweno_capi, only: HRW_ARITH_EXACT, HRW_ARITH_FAST, hrw_weno_calc_cnu, &
                          hrw_reconstruct_host, hrw_reconstruct_dev, hrw_message, &
                          hrw_reconstruct_lines_host, hrw_reconstruct_lines_dev
   implicit none
   private

   public :: weno, c1, c2, c3, HRW_ARITH_EXACT, HRW_ARITH_FAST

   !> Candidate-stencil coefficients c(j, r) of Shu (ICASE 97-65) Table 2.1, the same public
   !> tables as src/hrweno_weno.f90:15-21 (column r gives v at x_{i+1/2} from the stencil
   !> starting at cell i-r).
   real(rk), parameter :: c1(0:0, -1:0) = reshape([1.0_rk, 1.0_rk], [1, 2])
   real(rk), parameter :: c2(0:1, -1:1) = reshape([3.0_rk/2, -1.0_rk/2, &
                                                   1.0_rk/2, 1.0_rk/2, &
                                                   -1.0_rk/2, 3.0_rk/2], [2, 3])
   real(rk), parameter :: c3(0:2, -1:2) = reshape([11.0_rk/6, -7.0_rk/6, 1.0_rk/3, &
                                                   1.0_rk/3, 5.0_rk/6, -1.0_rk/6, &
                                                   -1.0_rk/6, 5.0_rk/6, 1.0_rk/3, &
                                                   1.0_rk/3, -7.0_rk/6, 11.0_rk/6], [3, 4])

   type :: weno
   !! WENO object (src/hrweno_weno.f90:23-46).
      character(:), allocatable :: msg
         !! error message
      integer :: ierr = 0
         !! error status
      integer :: ncells
         !! number of cells
      integer :: k = 3
         !! order of reconstruction within the cell (k = 1, 2 or 3)
      real(rk) :: eps = 1e-6_rk
         !! numerical smoothing factor
      integer :: arith = HRW_ARITH_FAST
         !! arithmetic form of the device kernels
      logical, private :: uniform_grid = .true.
         !! flag for uniform grid
      real(rk), allocatable :: cnu(:, :, :)
         !! cnu(0:k-1, -1:k-1, 1:ncells) for a non-uniform grid
   contains
      procedure, pass(self) :: reconstruct => weno_reconstruct
      procedure, pass(self) :: reconstruct_dev => weno_reconstruct_dev
      procedure, pass(self) :: reconstruct_lines => weno_reconstruct_lines
      procedure, pass(self) :: reconstruct_lines_dev => weno_reconstruct_lines_dev
   end type weno

   interface weno
      module procedure :: weno_init
   end interface weno

contains

   pure type(weno) function weno_init(ncells, k, eps, xedges, arith) result(self)
   !! weno(ncells, k, eps, xedges) (src/hrweno_weno.f90:54-127).  Same checks and messages;
   !! an absent k means k = 3 also for the cnu table.
      integer, intent(in) :: ncells
      integer, intent(in), optional :: k
      real(rk), intent(in), optional :: eps
      real(rk), intent(in), optional :: xedges(0:)
      integer, intent(in), optional :: arith
      integer(c_int) :: rc

      if (ncells > 0) then
         self%ncells = ncells
      else
         self%msg = "Invalid input 'ncells'. Valid range: ncells > 0."
         self%ierr = 1
         error stop self%msg
      end if

      if (present(k)) then
         if ((k >= 1) .and. (k <= 3)) then
            self%k = k
         else
            self%msg = "Invalid input 'k'. Valid range: 1 <= k <= 3."
            self%ierr = 1
            error stop self%msg
         end if
      end if

      if (present(eps)) then
         if (eps > epsilon(1.0_rk)) then
            self%eps = eps
         else
            self%msg = "Invalid input 'eps'. Valid range: eps > epsilon."
            self%ierr = 1
            error stop self%msg
         end if
      end if

      if (present(arith)) then
         if (arith == HRW_ARITH_EXACT .or. arith == HRW_ARITH_FAST) then
            self%arith = arith
         else
            self%msg = "Invalid input 'arith'. Valid values: HRW_ARITH_EXACT, HRW_ARITH_FAST."
            self%ierr = 1
            error stop self%msg
         end if
      end if

      if (present(xedges)) then
         if (size(xedges) == ncells + 1) then
            self%uniform_grid = .false.
            allocate (self%cnu(0:self%k - 1, -1:self%k - 1, 1:ncells))
            rc = hrw_weno_calc_cnu(ncells, self%k, xedges, self%cnu)
            if (rc /= 0) then
               self%msg = hrw_message()
               self%ierr = 1
               error stop self%msg
            end if
         else
            self%msg = "Invalid input 'xedges': size(xedges) /= ncells + 1."
            self%ierr = 1
            error stop self%msg
         end if
      end if

   end function weno_init

   pure subroutine weno_reconstruct(self, v, vl, vr)
   !! vl(i) = v_{i-1/2}^+ and vr(i) = v_{i+1/2}^- for every cell
   !! (src/hrweno_weno.f90:129-219), computed on the device.  v, vl and vr may be strided
   !! sections; they are passed contiguously.
      class(weno), intent(in), target :: self
      real(rk), intent(in) :: v(:)
      real(rk), intent(out) :: vl(:)
      real(rk), intent(out) :: vr(:)
      type(c_ptr) :: pc
      integer(c_int) :: rc

      if (size(v) /= self%ncells .or. size(vl) /= self%ncells .or. size(vr) /= self%ncells) &
         error stop "Invalid input 'v', 'vl' or 'vr': size /= ncells."
      pc = c_null_ptr
      if (.not. self%uniform_grid) pc = c_loc(self%cnu)
      rc = hrw_reconstruct_host(self%ncells, self%k, self%eps, self%arith, pc, v, 1_c_int64_t, &
                                vl, 1_c_int64_t, vr, 1_c_int64_t)
      if (rc /= 0) error stop hrw_message()

   end subroutine weno_reconstruct

   subroutine weno_reconstruct_dev(self, v, vl, vr, incv, incl, incr, stream)
   !! The same on device arrays (type(c_ptr)), uniform grids, optional element strides;
   !! enqueued on `stream` (default: the null stream).
      class(weno), intent(in) :: self
      type(c_ptr), intent(in) :: v, vl, vr
      integer, intent(in), optional :: incv, incl, incr
      type(c_ptr), intent(in), optional :: stream
      integer(c_int64_t) :: iv, il, ir
      type(c_ptr) :: s
      integer(c_int) :: rc

      if (.not. self%uniform_grid) error stop "reconstruct_dev: non-uniform grids use reconstruct."
      iv = 1; il = 1; ir = 1
      if (present(incv)) iv = incv
      if (present(incl)) il = incl
      if (present(incr)) ir = incr
      s = c_null_ptr
      if (present(stream)) s = stream
      rc = hrw_reconstruct_dev(self%ncells, self%k, self%eps, self%arith, v, iv, vl, il, vr, ir, s)
      if (rc /= 0) error stop hrw_message()

   end subroutine weno_reconstruct_dev

   subroutine weno_reconstruct_lines(self, v, vl, vr, nlines, incv, ldv, inco, ldo)
   !! %reconstruct on nlines lines of one array in one device launch: element i of line l
   !! (1-based) is v(1 + (l-1)*ldv + (i-1)*incv), and likewise vl, vr with inco, ldo.  The
   !! x1 lines of example 2's state u((j-1)*nc1 + i) are nlines = nc2, incv = 1, ldv = nc1;
   !! its x2 lines nlines = nc1, incv = nc1, ldv = 1 (example/example2_pbe_2d_fv.f90:95-112).
   !! Elements of vl, vr between the lines keep their values.
      class(weno), intent(in), target :: self
      real(rk), intent(in) :: v(:)
      real(rk), intent(inout) :: vl(:), vr(:)
      integer, intent(in) :: nlines, incv, ldv, inco, ldo
      type(c_ptr) :: pc
      integer(c_int) :: rc
      integer(c_int64_t) :: nv, no

      if (nlines < 1 .or. incv < 1 .or. inco < 1 .or. ldv < 0 .or. ldo < 0) &
         error stop "Invalid input 'nlines' or strides."
      nv = int(nlines - 1, c_int64_t)*ldv + int(self%ncells - 1, c_int64_t)*incv + 1
      no = int(nlines - 1, c_int64_t)*ldo + int(self%ncells - 1, c_int64_t)*inco + 1
      if (size(v, kind=c_int64_t) < nv .or. size(vl, kind=c_int64_t) < no .or. &
          size(vr, kind=c_int64_t) < no) error stop "Invalid input 'v', 'vl' or 'vr': too small."
      pc = c_null_ptr
      if (.not. self%uniform_grid) pc = c_loc(self%cnu)
      rc = hrw_reconstruct_lines_host(self%ncells, self%k, self%eps, self%arith, pc, &
                                      int(nlines, c_int64_t), v, int(incv, c_int64_t), &
                                      int(ldv, c_int64_t), vl, vr, int(inco, c_int64_t), &
                                      int(ldo, c_int64_t))
      if (rc /= 0) error stop hrw_message()
   end subroutine weno_reconstruct_lines

   subroutine weno_reconstruct_lines_dev(self, v, vl, vr, nlines, incv, ldv, inco, ldo, stream)
   !! The same on device arrays (uniform grids), enqueued on `stream` (default: null stream).
      class(weno), intent(in) :: self
      type(c_ptr), intent(in) :: v, vl, vr
      integer, intent(in) :: nlines, incv, ldv, inco, ldo
      type(c_ptr), intent(in), optional :: stream
      type(c_ptr) :: s
      integer(c_int) :: rc
      if (.not. self%uniform_grid) error stop "reconstruct_lines_dev: non-uniform grids use reconstruct_lines."
      s = c_null_ptr
      if (present(stream)) s = stream
      rc = hrw_reconstruct_lines_dev(self%ncells, self%k, self%eps, self%arith, &
                                     int(nlines, c_int64_t), v, int(incv, c_int64_t), &
                                     int(ldv, c_int64_t), vl, vr, int(inco, c_int64_t), &
                                     int(ldo, c_int64_t), s)
      if (rc /= 0) error stop hrw_message()
   end subroutine weno_reconstruct_lines_dev

end module hrweno_weno
