!> Working precision of the host-side API (reference: src/precision.F90:6-15).
!> The engine's host arrays are FP64; reduced-precision storage is selected per grid
!> (alloc_grid's `storage` argument), not by recompiling, so `wp` is always double.
module precision

   implicit none
   private

   public :: wp

   integer, parameter :: dp = kind(1.0d0)
   integer, parameter :: wp = dp

end module
