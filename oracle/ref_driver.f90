! Test infrastructure only: a bind(C) driver over the UNMODIFIED reference module
! API (ivan-pi/periodic-lbm, src/), compiled together with the reference sources by
! oracle/Makefile into oracle/_ref/libplbm_ref.so.  It is the checker that pins the
! C restatement in oracle/oracle.c and produces the golden fixtures; nothing in the
! product links or loads it.
!
! Every entry point forwards to the reference routine it names:
!   alloc_grid / dealloc_grid / set_properties / set_pdf_to_equilibrium /
!   update_macros / equilibrium          src/fvm_bardow.F90:99-390
!   perform_lbm_step / lbm_stream        src/periodic_lbm.f90:15-129
!   collide_bgk                          src/collision_bgk.F90:17-82
!   collide_trt, magic_number, lambda_d  src/collision_trt.F90:24-160
!   collide_rr                           src/collision_regularized.F90:21-202
!   taylor_green_t%eval                  src/benchmarks/taylor_green.f90:31-84
!   shear_layer_t eval_shear_layer_grid  src/benchmarks/shear_layer.f90:25-52
!   perform_triple_step                  src/fvm_bardow.F90:322-340
!   perform_dugks_step, dugks_collide,
!   dugks_stream                         src/periodic_dugks.F90:25-299
module ref_driver

   use iso_c_binding
   use fvm_bardow
   use periodic_lbm, only: perform_lbm_step, lbm_stream
   use collision_bgk, only: collide_bgk
   use collision_trt, only: collide_trt, magic_number, lambda_d
   use collision_regularized, only: collide_rr
   use taylor_green, only: taylor_green_t
   use shear_layer_case, only: shear_layer_t
   use barotropic_vortex_case, only: vortex_case_t
   use periodic_dugks, only: perform_dugks_step, dugks_collide, dugks_stream

   implicit none

   type :: grid_box
      type(lattice_grid) :: grid
   end type

contains

   function unbox(h) result(p)
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      call c_f_pointer(h, p)
   end function

   subroutine select_collision(grid, model)
      type(lattice_grid), intent(inout) :: grid
      integer(c_int), intent(in) :: model
      select case (model)
      case (0)
         grid%collision => collide_bgk
      case (1)
         grid%collision => collide_trt
      case (2)
         grid%collision => collide_rr
      case (3)
         ! collide_bgk as compiled: bgk_kernel_cache in the -DSPLIT build (ref-split)
         grid%collision => collide_bgk
      case default
         error stop "ref_driver: unknown collision model"
      end select
   end subroutine

   function ref_create(nx, ny, nf, model) result(h) bind(c, name="ref_create")
      integer(c_int), value :: nx, ny, nf, model
      type(c_ptr) :: h
      type(grid_box), pointer :: p
      allocate(p)
      call alloc_grid(p%grid, nx, ny, nf=nf, log=.false.)
      p%grid%streaming => lbm_stream
      call select_collision(p%grid, model)
      ! the reference leaves f uninitialised; zero it so dumps are defined
      p%grid%f = 0.0_wp
      h = c_loc(p)
   end function

   subroutine ref_destroy(h) bind(c, name="ref_destroy")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call dealloc_grid(p%grid)
      deallocate(p)
   end subroutine

   subroutine ref_set_model(h, model) bind(c, name="ref_set_model")
      type(c_ptr), value :: h
      integer(c_int), value :: model
      type(grid_box), pointer :: p
      p => unbox(h)
      call select_collision(p%grid, model)
   end subroutine

   subroutine ref_set_properties(h, nu, dt, magic, has_magic) bind(c, name="ref_set_properties")
      type(c_ptr), value :: h
      real(c_double), value :: nu, dt, magic
      integer(c_int), value :: has_magic
      type(grid_box), pointer :: p
      p => unbox(h)
      if (has_magic /= 0) then
         call set_properties(p%grid, nu, dt, magic)
      else
         call set_properties(p%grid, nu, dt)
      end if
   end subroutine

   subroutine ref_set_omega(h, omega, magic) bind(c, name="ref_set_omega")
      ! the two collision parameters set_properties derives, assigned directly (no
      ! "trt magic" print on stdout, for callers that own stdout)
      type(c_ptr), value :: h
      real(c_double), value :: omega, magic
      type(grid_box), pointer :: p
      p => unbox(h)
      p%grid%omega = omega
      p%grid%trt_magic = magic
   end subroutine

   subroutine ref_get_params(h, out) bind(c, name="ref_get_params")
      ! out = [nu, dt, tau, omega, trt_magic, csqr]
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(6)
      type(grid_box), pointer :: p
      p => unbox(h)
      out = [p%grid%nu, p%grid%dt, p%grid%tau, p%grid%omega, p%grid%trt_magic, p%grid%csqr]
   end subroutine

   subroutine ref_set_macros(h, rho, ux, uy) bind(c, name="ref_set_macros")
      type(c_ptr), value :: h
      real(c_double), intent(in) :: rho(*), ux(*), uy(*)
      type(grid_box), pointer :: p
      integer :: n
      p => unbox(h)
      n = p%grid%nx * p%grid%ny
      p%grid%rho = reshape(rho(1:n), [p%grid%ny, p%grid%nx])
      p%grid%ux  = reshape(ux(1:n),  [p%grid%ny, p%grid%nx])
      p%grid%uy  = reshape(uy(1:n),  [p%grid%ny, p%grid%nx])
   end subroutine

   subroutine ref_get_macros(h, rho, ux, uy) bind(c, name="ref_get_macros")
      type(c_ptr), value :: h
      real(c_double), intent(out) :: rho(*), ux(*), uy(*)
      type(grid_box), pointer :: p
      integer :: n
      p => unbox(h)
      n = p%grid%nx * p%grid%ny
      rho(1:n) = reshape(p%grid%rho, [n])
      ux(1:n)  = reshape(p%grid%ux,  [n])
      uy(1:n)  = reshape(p%grid%uy,  [n])
   end subroutine

   subroutine ref_set_pdf_to_equilibrium(h) bind(c, name="ref_set_pdf_to_equilibrium")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call set_pdf_to_equilibrium(p%grid)
   end subroutine

   subroutine ref_perform_lbm_step(h, nsteps) bind(c, name="ref_perform_lbm_step")
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      type(grid_box), pointer :: p
      integer :: s
      p => unbox(h)
      do s = 1, nsteps
         call perform_lbm_step(p%grid)
      end do
   end subroutine

   subroutine ref_stream(h) bind(c, name="ref_stream")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call lbm_stream(p%grid)
   end subroutine

   subroutine ref_set_streaming(h, op) bind(c, name="ref_set_streaming")
      ! 0 lbm_stream, 1 stream_fvm_bardow, 2 stream_fdm_bardow, 3 stream_fdm_sofonea
      type(c_ptr), value :: h
      integer(c_int), value :: op
      type(grid_box), pointer :: p
      p => unbox(h)
      select case (op)
      case (0)
         p%grid%streaming => lbm_stream
      case (1)
         p%grid%streaming => stream_fvm_bardow
      case (2)
         p%grid%streaming => stream_fdm_bardow
      case (3)
         p%grid%streaming => stream_fdm_sofonea
      case default
         error stop "ref_driver: unknown streaming operator"
      end select
   end subroutine

   subroutine ref_streaming(h) bind(c, name="ref_streaming")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call p%grid%streaming()
   end subroutine

   subroutine ref_perform_step(h, nsteps) bind(c, name="ref_perform_step")
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      type(grid_box), pointer :: p
      integer :: s
      p => unbox(h)
      do s = 1, nsteps
         call perform_step(p%grid)
      end do
   end subroutine

   subroutine ref_perform_triple_step(h, nsteps) bind(c, name="ref_perform_triple_step")
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      type(grid_box), pointer :: p
      integer :: s
      p => unbox(h)
      do s = 1, nsteps
         call perform_triple_step(p%grid)
      end do
   end subroutine

   subroutine ref_set_dugks(h) bind(c, name="ref_set_dugks")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      p%grid%collision => dugks_collide
      p%grid%streaming => dugks_stream
   end subroutine

   subroutine ref_perform_dugks_step(h, nsteps) bind(c, name="ref_perform_dugks_step")
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
      type(grid_box), pointer :: p
      integer :: s
      p => unbox(h)
      do s = 1, nsteps
         call perform_dugks_step(p%grid)
      end do
   end subroutine

   subroutine ref_collide(h) bind(c, name="ref_collide")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call p%grid%collision()
   end subroutine

   subroutine ref_update_macros(h) bind(c, name="ref_update_macros")
      type(c_ptr), value :: h
      type(grid_box), pointer :: p
      p => unbox(h)
      call update_macros(p%grid)
   end subroutine

   subroutine ref_get_indices(h, idx) bind(c, name="ref_get_indices")
      ! idx = [iold, inew, imid, ld] (1-based buffer indices as in the reference)
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: idx(4)
      type(grid_box), pointer :: p
      p => unbox(h)
      idx = [p%grid%iold, p%grid%inew, p%grid%imid, size(p%grid%f, 1)]
   end subroutine

   subroutine ref_get_f(h, buf, f) bind(c, name="ref_get_f")
      ! f(ld,nx,0:8) of buffer `buf` (1-based), Fortran order
      type(c_ptr), value :: h
      integer(c_int), value :: buf
      real(c_double), intent(out) :: f(*)
      type(grid_box), pointer :: p
      integer :: n
      p => unbox(h)
      n = size(p%grid%f, 1) * p%grid%nx * 9
      f(1:n) = reshape(p%grid%f(:,:,:,buf), [n])
   end subroutine

   subroutine ref_set_f(h, buf, f) bind(c, name="ref_set_f")
      type(c_ptr), value :: h
      integer(c_int), value :: buf
      real(c_double), intent(in) :: f(*)
      type(grid_box), pointer :: p
      integer :: n
      p => unbox(h)
      n = size(p%grid%f, 1) * p%grid%nx * 9
      p%grid%f(:,:,:,buf) = reshape(f(1:n), [size(p%grid%f, 1), p%grid%nx, 9])
   end subroutine

   subroutine ref_equilibrium(rho, ux, uy, feq) bind(c, name="ref_equilibrium")
      real(c_double), value :: rho, ux, uy
      real(c_double), intent(out) :: feq(0:8)
      feq = equilibrium(rho, ux, uy)
   end subroutine

   function ref_magic_number(le, ld) result(r) bind(c, name="ref_magic_number")
      real(c_double), value :: le, ld
      real(c_double) :: r
      r = magic_number(le, ld)
   end function

   function ref_lambda_d(omega, x) result(r) bind(c, name="ref_lambda_d")
      real(c_double), value :: omega, x
      real(c_double) :: r
      r = lambda_d(omega, x)
   end function

   subroutine ref_tgv_eval(nx, ny, kx, ky, umax, nu, t, p, ux, uy) bind(c, name="ref_tgv_eval")
      integer(c_int), value :: nx, ny
      real(c_double), value :: kx, ky, umax, nu, t
      real(c_double), intent(out) :: p(ny,nx), ux(ny,nx), uy(ny,nx)
      type(taylor_green_t) :: tg
      tg = taylor_green_t(nx, ny, kx, ky, umax, nu)
      call tg%eval(t, p, ux, uy)
   end subroutine

   subroutine ref_vortex_eval(nx, ny, u0, xc, yc, rc, eps, rho, ux, uy) bind(c, name="ref_vortex_eval")
      integer(c_int), value :: nx, ny
      real(c_double), value :: u0, xc, yc, rc, eps
      real(c_double), intent(out) :: rho(ny,nx), ux(ny,nx), uy(ny,nx)
      type(vortex_case_t) :: vc
      vc = vortex_case_t(u0, xc, yc, rc, eps=eps)
      call vc%eval(nx, ny, rho, ux, uy)
   end subroutine

   subroutine ref_shear_eval(nx, ny, length, u0, k, delta, rho, ux, uy) bind(c, name="ref_shear_eval")
      integer(c_int), value :: nx, ny
      real(c_double), value :: length, u0, k, delta
      real(c_double), intent(out) :: rho(ny,nx), ux(ny,nx), uy(ny,nx)
      type(shear_layer_t) :: sl
      sl%length = length
      sl%U0 = u0
      sl%k = k
      sl%delta = delta
      ! `eval` binds eval_shear_layer_unstructured; the grid form is its own binding
      ! (src/benchmarks/shear_layer.f90:17-18)
      call sl%eval_shear_layer_grid(nx, ny, rho, ux, uy)
   end subroutine

end module
