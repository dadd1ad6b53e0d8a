! Declarations of the two Macaw modules the cmf2_base drop-in reads, with the
! attributes the reference gives them (dvr.f90:11-26, hamilton.f90:24-49:
! allocatable module arrays without TARGET).  Only the names the drop-in
! uses; the rest of those modules (their procedures) does not compile with
! flang without changes to the reference (GNU COMPLEX(x, y) intrinsic), so the
! compile test of the drop-in uses these declarations in their place.
module dvr
   implicit none
   private
   public :: dvrType, ngp, ngpMax, grid, d1DVR, xd1DVR, d2DVR, gridOper
   character(6), dimension(:), allocatable :: dvrType
   integer :: ngpMax
   integer, dimension(:), allocatable :: ngp
   double precision, dimension(:,:), allocatable :: grid
   double precision, dimension(:,:,:), allocatable :: d1DVR, xd1DVR, d2DVR
   double precision, dimension(:,:,:), allocatable :: gridOper
end module dvr

module hamilton
   implicit none
   private
   public :: nHam, ae, iOpDof, lhs_el, rhs_el, nOpParam
   integer :: nHam
   integer :: nOpParam
   integer, dimension(:), allocatable :: lhs_el, rhs_el
   integer, dimension(:,:), allocatable :: iOpDof
   double complex, dimension(:), allocatable :: ae
end module hamilton
