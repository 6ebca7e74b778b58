!> Exact periodic streaming and the stream-collide step, under the reference's module
!> name (ivan-pi/periodic-lbm src/periodic_lbm.f90:9-11).  Both run on the GPU engine:
!> lbm_stream is the device pull-streaming operator, and perform_lbm_step launches one
!> fused pull-stream + collide kernel when grid%streaming is lbm_stream and
!> grid%collision is one of collide_bgk / collide_trt / collide_rr.
module periodic_lbm

   use precision, only: wp
   use fvm_bardow, only: lattice_grid, perform_step, lbm_stream => plbm_lbm_stream

   implicit none
   private

   public :: perform_lbm_step
   public :: lbm_stream

contains

   !> perform_lbm_step (src/periodic_lbm.f90:15-29): stream iold -> inew, collide inew,
   !> swap iold and inew.
   subroutine perform_lbm_step(grid)
      type(lattice_grid), intent(inout) :: grid
      call perform_step(grid)
   end subroutine

end module
