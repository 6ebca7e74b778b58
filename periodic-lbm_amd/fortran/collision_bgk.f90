!> BGK collision on the GPU engine, under the reference's module name
!> (ivan-pi/periodic-lbm src/collision_bgk.F90:9, collide_bgk :17-82).
module collision_bgk

   use fvm_bardow, only: collide_bgk => plbm_collide_bgk

   implicit none
   private

   public :: collide_bgk

end module
