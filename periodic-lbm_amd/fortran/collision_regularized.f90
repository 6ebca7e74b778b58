!> Recursive-regularized collision on the GPU engine, under the reference's module name
!> (ivan-pi/periodic-lbm src/collision_regularized.F90:9, collide_rr :21-202).
module collision_regularized

   use fvm_bardow, only: collide_rr => plbm_collide_rr

   implicit none
   private

   public :: collide_rr

end module
