!> Convected barotropic vortex initial field, under the reference's module name
!> (ivan-pi/periodic-lbm src/benchmarks/barotropic_vortex_case.F90; Wissocq, Boussuge &
!> Sagaut 2020).  The grid form is evaluated by the engine's host routine
!> plbm_field_barotropic_vortex.
module barotropic_vortex_case

   use iso_c_binding
   use precision, only: wp

   implicit none
   private

   public :: vortex_case_t

   type :: vortex_case_t
      real(wp) :: U0          !> Advection velocity
      real(wp) :: xc, yc      !> Position of vortex center
      real(wp) :: Rc          !> Characteristic radius
      real(wp) :: eps         !> Strength of vortex
      real(wp) :: rho0 = 1.0_wp
      real(wp) :: csqr = 1._wp/3._wp
   contains
      procedure :: eval => eval_vortex_case, &
                           eval_vortex_case_unstructured
   end type

   interface
      integer(c_int) function plbm_field_barotropic_vortex(nx, ny, u0, xc, yc, rc, eps, rho0, &
            csqr, rho, ux, uy) bind(c, name="plbm_field_barotropic_vortex")
         import :: c_int, c_double
         integer(c_int), value :: nx, ny
         real(c_double), value :: u0, xc, yc, rc, eps, rho0, csqr
         real(c_double), intent(out) :: rho(*), ux(*), uy(*)
      end function
   end interface

contains

   !> eval_vortex_case (src/benchmarks/barotropic_vortex_case.F90:34-83) on the (ny, nx)
   !> cell centres.
   subroutine eval_vortex_case(case,nx,ny,rho,ux,uy)
      class(vortex_case_t), intent(in) :: case
      integer, intent(in) :: nx, ny
      real(wp), intent(out), contiguous :: rho(:,:)
      real(wp), intent(out), contiguous :: ux(:,:), uy(:,:)
      if (plbm_field_barotropic_vortex(int(nx,c_int), int(ny,c_int), case%U0, case%xc, &
            case%yc, case%Rc, case%eps, case%rho0, case%csqr, rho, ux, uy) /= 0) then
         error stop "vortex_case_t%eval: invalid case parameters"
      end if
   end subroutine

   !> eval_vortex_case_unstructured (src/benchmarks/barotropic_vortex_case.F90:85-130)
   subroutine eval_vortex_case_unstructured(case,n,xy,rho,ux,uy)
      class(vortex_case_t), intent(in) :: case
      integer, intent(in) :: n
      real(wp), intent(in) :: xy(:,:)
      real(wp), intent(out) :: rho(:)
      real(wp), intent(out) :: ux(:), uy(:)
      real(wp) :: xr, yr, rsqr, g
      integer :: i
      do i = 1, n
         xr = xy(1,i) - case%xc
         yr = xy(2,i) - case%yc
         rsqr = xr*xr + yr*yr
         g = exp(-0.5_wp*rsqr/case%Rc**2)
         ux(i) = case%U0 - case%eps*(yr/case%Rc)*g
         uy(i) = case%eps*(xr/case%Rc)*g
         rho(i) = case%rho0*exp(-0.5_wp*(case%eps**2/case%csqr)*exp(-rsqr/case%Rc**2))
      end do
   end subroutine

end module
