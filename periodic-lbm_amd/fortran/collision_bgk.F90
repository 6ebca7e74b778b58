!> BGK collision on the GPU engine, under the reference's module name
!> (ivan-pi/periodic-lbm src/collision_bgk.F90:9, collide_bgk :17-82).  As in the
!> reference, compiling with -DSPLIT selects the split-weight algebra
!> (bgk_kernel_cache, src/collision_bgk.F90:23,84-176).
module collision_bgk

#if SPLIT
   use fvm_bardow, only: collide_bgk => plbm_collide_bgk_split
#else
   use fvm_bardow, only: collide_bgk => plbm_collide_bgk
#endif

   implicit none
   private

   public :: collide_bgk

end module
