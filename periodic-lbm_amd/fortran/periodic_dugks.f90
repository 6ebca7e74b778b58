!> DUGKS on the GPU engine, under the reference's module name and API
!> (ivan-pi/periodic-lbm src/periodic_dugks.F90:1-10: perform_dugks_step, dugks_stream,
!> dugks_collide).  The operators live in the fvm_bardow shim so that its dispatchers know
!> them as device operators.
module periodic_dugks

   use fvm_bardow, only: perform_dugks_step => plbm_grid_perform_dugks_step, &
                         dugks_stream => plbm_grid_dugks_stream, &
                         dugks_collide => plbm_grid_dugks_collide

   implicit none
   private

   public :: perform_dugks_step
   public :: dugks_stream
   public :: dugks_collide

end module
