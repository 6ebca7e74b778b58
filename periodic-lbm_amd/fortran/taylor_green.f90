!> Analytic decaying Taylor-Green vortex, under the reference's module name
!> (ivan-pi/periodic-lbm src/benchmarks/taylor_green.f90).  The velocity and pressure
!> fields are evaluated by the engine's host routine plbm_field_taylor_green (same
!> formula and evaluation order); the optional strain rate S is evaluated here.
module taylor_green

   use iso_c_binding
   use precision, only: wp
   use fvm_bardow, only: lattice_grid

   implicit none
   private

   public :: taylor_green_t
   public :: pi

   type :: taylor_green_t
      integer :: nx, ny
      real(wp) :: kx, ky
      real(wp) :: umax, nu
      real(wp) :: td
   contains
      procedure :: eval => taylor_green_eval
      procedure :: decay_time => taylor_green_decay_time
   end type

   real(wp), parameter :: pi = 4._wp*atan(1._wp)

   interface taylor_green_t
      module procedure :: taylor_green_t_constructor
   end interface

   interface
      integer(c_int) function plbm_field_taylor_green(nx, ny, kx, ky, umax, nu, t, p, ux, uy) &
            bind(c, name="plbm_field_taylor_green")
         import :: c_int, c_double
         integer(c_int), value :: nx, ny
         real(c_double), value :: kx, ky, umax, nu, t
         real(c_double), intent(out) :: p(*), ux(*), uy(*)
      end function
   end interface

contains

   !> Constructor (src/benchmarks/taylor_green.f90:31-43): decay time 1/(nu (kx^2+ky^2)).
   function taylor_green_t_constructor(nx,ny,kx,ky,umax,nu) result(this)
      integer, intent(in) :: nx, ny
      real(wp), intent(in) :: kx, ky, umax, nu
      type(taylor_green_t) :: this
      this%nx = nx
      this%ny = ny
      this%kx = kx
      this%ky = ky
      this%umax = umax
      this%nu = nu
      this%td = 1._wp/(nu*(kx**2 + ky**2))
   end function

   function taylor_green_decay_time(self) result(tc)
      class(taylor_green_t), intent(in) :: self
      real(wp) :: tc
      tc = self%td
   end function

   !> eval (src/benchmarks/taylor_green.f90:51-84) at time t on the (ny, nx) cell centres.
   subroutine taylor_green_eval(self,t,p,ux,uy,S)
      class(taylor_green_t), intent(in) :: self
      real(wp), intent(in) :: t
      real(wp), intent(out), contiguous :: p(:,:), ux(:,:), uy(:,:)
      real(wp), intent(out), optional :: S(self%ny,self%nx,3)

      integer :: x, y
      real(wp) :: xx, yy, e

      if (plbm_field_taylor_green(int(self%nx,c_int), int(self%ny,c_int), self%kx, self%ky, &
            self%umax, self%nu, t, p, ux, uy) /= 0) then
         error stop "taylor_green_t%eval: invalid case parameters"
      end if

      if (present(S)) then
         associate(umax => self%umax, kx => self%kx, ky => self%ky)
         e = exp(-t/self%td)
         do x = 1, self%nx
            xx = (x - 1) + 0.5_wp
            do y = 1, self%ny
               yy = (y - 1) + 0.5_wp
               S(y,x,1) = umax*sqrt(kx*ky)*sin(kx*xx)*sin(ky*yy)*e
               S(y,x,2) = 0.5_wp*umax*(sqrt(kx**3/ky) - sqrt(ky**3/kx))*cos(kx*xx)*cos(ky*yy)*e
               S(y,x,3) = -S(y,x,1)
            end do
         end do
         end associate
      end if
   end subroutine

end module
