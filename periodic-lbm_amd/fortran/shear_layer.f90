!> Doubly periodic shear layer initial field, under the reference's module name
!> (ivan-pi/periodic-lbm src/benchmarks/shear_layer.f90).  The grid form is evaluated by
!> the engine's host routine plbm_field_shear_layer.
module shear_layer_case

   use iso_c_binding
   use precision, only: wp

   implicit none
   private

   public :: shear_layer_t

   type :: shear_layer_t
      real(wp) :: length
      real(wp) :: U0
      real(wp) :: k
      real(wp) :: delta
      real(wp) :: rho0 = 1.0_wp
   contains
      procedure :: eval => eval_shear_layer_unstructured, &
                           eval_shear_layer_grid
   end type

   real(wp), parameter :: pi = 4.0_wp*atan(1.0_wp)

   interface
      integer(c_int) function plbm_field_shear_layer(nx, ny, length, u0, k, delta, rho, ux, uy) &
            bind(c, name="plbm_field_shear_layer")
         import :: c_int, c_double
         integer(c_int), value :: nx, ny
         real(c_double), value :: length, u0, k, delta
         real(c_double), intent(out) :: rho(*), ux(*), uy(*)
      end function
   end interface

contains

   !> eval_shear_layer_grid (src/benchmarks/shear_layer.f90:25-52)
   subroutine eval_shear_layer_grid(case,nx,ny,rho,ux,uy)
      class(shear_layer_t), intent(in) :: case
      integer, intent(in) :: nx, ny
      real(wp), intent(out), contiguous :: rho(:,:), ux(:,:), uy(:,:)
      if (plbm_field_shear_layer(int(nx,c_int), int(ny,c_int), case%length, case%U0, &
            case%k, case%delta, rho, ux, uy) /= 0) then
         error stop "shear_layer_t%eval: invalid case parameters"
      end if
      rho = case%rho0
   end subroutine

   !> eval_shear_layer_unstructured (src/benchmarks/shear_layer.f90:54-81): points
   !> xy(2,n) in absolute coordinates.
   subroutine eval_shear_layer_unstructured(case,n,xy,rho,ux,uy)
      class(shear_layer_t), intent(in) :: case
      integer, intent(in) :: n
      real(wp), intent(in) :: xy(2,n)
      real(wp), intent(out) :: rho(n), ux(n), uy(n)
      real(wp) :: xd, yd
      integer :: i
      do i = 1, n
         xd = xy(1,i)/case%length
         yd = xy(2,i)/case%length
         uy(i) = case%U0 * case%delta * sin(2*pi*(xd + 0.25_wp))
         if (yd <= 0.5_wp) then
            ux(i) = case%U0 * tanh(case%k * (yd - 0.25_wp))
         else
            ux(i) = case%U0 * tanh(case%k * (0.75_wp - yd))
         end if
         rho(i) = case%rho0
      end do
   end subroutine

end module
