!> Two-relaxation-time collision on the GPU engine, under the reference's module name
!> (ivan-pi/periodic-lbm src/collision_trt.F90:9-10).
module collision_trt

   use precision, only: wp
   use fvm_bardow, only: collide_trt => plbm_collide_trt

   implicit none
   private

   public :: magic_number, lambda_d
   public :: collide_trt

contains

   !> Magic parameter of the TRT pair of rates (src/collision_trt.F90:24-28)
   pure real(wp) function magic_number(le,ld)
      real(wp), intent(in) :: le, ld
      magic_number = (2.0_wp - le) * (2.0_wp - ld) / (4.0_wp*le*ld)
   end function

   !> Odd rate for a given even rate and magic parameter (src/collision_trt.F90:30-34)
   pure real(wp) function lambda_d(omega,x)
      real(wp), intent(in) :: omega, x
      lambda_d = (4.0_wp - 2.0_wp*omega) / (4.0_wp * x * omega + 2.0_wp - omega)
   end function

end module
