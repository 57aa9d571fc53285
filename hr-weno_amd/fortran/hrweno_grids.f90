module hrweno_grids
!! 1D grid class with the public API of the reference module hrweno_grids
!! (src/hrweno_grids.f90): grid1 with %linear, %bilinear, %log and %geometric.  The edges,
!! centres and widths are computed by libhrweno_hip's host routines (the arithmetic of
!! src/hrweno_grids.f90:76-78, 122-131, 171-175, 222-225, 246-247) so the Fortran, C and
!! Python hosts share one implementation; errors stop with the reference's messages.
   use, intrinsic :: iso_c_binding, only: c_int
   use hrweno_kinds, only: rk
   use hrweno_capi, only: hrw_grid_linear, hrw_grid_bilinear, hrw_grid_log, &
                          hrw_grid_geometric, hrw_grid_compute, hrw_message
   implicit none
   private

   public :: grid1

   type :: grid1
   !! 1D grid (src/hrweno_grids.f90:12-37).
      character(:), allocatable :: name
         !! variable name
      character(:), allocatable :: scale
         !! scale type
      integer :: ncells
         !! number of cells
      real(rk), allocatable :: edges(:)
         !! vector(0:ncells) of cell edges
      real(rk), allocatable :: center(:)
         !! vector(ncells) of cell centers
      real(rk), allocatable :: width(:)
         !! vector(ncells) of cell widths
      real(rk), dimension(:), pointer :: left => null()
         !! vector(ncells) of left cell boundaries
      real(rk), dimension(:), pointer :: right => null()
         !! vector(ncells) of right cell boundaries
   contains
      procedure, pass(self) :: linear => grid1_linear
      procedure, pass(self) :: bilinear => grid1_bilinear
      procedure, pass(self) :: log => grid1_log
      procedure, pass(self) :: geometric => grid1_geometric
      procedure, pass(self), private :: clear => grid1_clear
      procedure, pass(self), private :: compute => grid1_compute
   end type grid1

contains

   pure subroutine grid1_linear(self, xmin, xmax, ncells, name)
   !! Linear grid: edges xmin + (xmax-xmin)/ncells*i.
      class(grid1), intent(inout), target :: self
      real(rk), intent(in) :: xmin, xmax
      integer, intent(in) :: ncells
      character(*), intent(in), optional :: name
      real(rk) :: xedges(0:max(ncells, 0))
      if (hrw_grid_linear(xmin, xmax, int(ncells, c_int), xedges) /= 0) error stop hrw_message()
      call self%clear
      self%scale = "linear"
      call self%compute(xedges, name)
   end subroutine grid1_linear

   pure subroutine grid1_bilinear(self, xmin, xcross, xmax, ncells, name)
   !! Two linear grids in series, [xmin, xcross] and [xcross, xmax].
      class(grid1), intent(inout), target :: self
      real(rk), intent(in) :: xmin, xcross, xmax
      integer, intent(in) :: ncells(2)
      character(*), intent(in), optional :: name
      real(rk) :: xedges(0:max(sum(ncells), 0))
      if (hrw_grid_bilinear(xmin, xcross, xmax, int(ncells(1), c_int), int(ncells(2), c_int), &
                            xedges) /= 0) error stop hrw_message()
      call self%clear
      self%scale = "bilinear"
      call self%compute(xedges, name)
   end subroutine grid1_bilinear

   pure subroutine grid1_log(self, xmin, xmax, ncells, name)
   !! Linear grid in y = log(x).
      class(grid1), intent(inout), target :: self
      real(rk), intent(in) :: xmin, xmax
      integer, intent(in) :: ncells
      character(*), intent(in), optional :: name
      real(rk) :: xedges(0:max(ncells, 0))
      if (hrw_grid_log(xmin, xmax, int(ncells, c_int), xedges) /= 0) error stop hrw_message()
      call self%clear
      self%scale = "log"
      call self%compute(xedges, name)
   end subroutine grid1_log

   subroutine grid1_geometric(self, xmin, xmax, ratio, ncells, name)
   !! Geometric grid, width(i+1) = ratio*width(i).
      class(grid1), intent(inout), target :: self
      real(rk), intent(in) :: xmin, xmax, ratio
      integer, intent(in) :: ncells
      character(*), intent(in), optional :: name
      real(rk) :: xedges(0:max(ncells, 0))
      if (hrw_grid_geometric(xmin, xmax, ratio, int(ncells, c_int), xedges) /= 0) &
         error stop hrw_message()
      call self%clear
      self%scale = "geometric"
      call self%compute(xedges, name)
   end subroutine grid1_geometric

   pure subroutine grid1_compute(self, xedges, name)
   !! Map edges to the object: left/right views, centres and widths.
      class(grid1), intent(inout), target :: self
      real(rk), intent(in) :: xedges(0:)
      character(*), intent(in), optional :: name
      integer :: n
      n = ubound(xedges, 1)
      self%ncells = n
      self%edges = xedges
      self%left => self%edges(0:n - 1)
      self%right => self%edges(1:n)
      allocate (self%center(n), self%width(n))
      if (hrw_grid_compute(int(n, c_int), xedges, self%center, self%width) /= 0) &
         error stop hrw_message()
      if (present(name)) then
         self%name = name
      else
         self%name = ""
      end if
   end subroutine grid1_compute

   pure subroutine grid1_clear(self)
   !! Reset the object.
      class(grid1), intent(inout), target :: self
      if (allocated(self%name)) deallocate (self%name)
      if (allocated(self%edges)) deallocate (self%edges)
      if (allocated(self%center)) deallocate (self%center)
      if (allocated(self%width)) deallocate (self%width)
      nullify (self%left, self%right)
      self%ncells = 0
      self%scale = ""
   end subroutine grid1_clear

end module hrweno_grids
