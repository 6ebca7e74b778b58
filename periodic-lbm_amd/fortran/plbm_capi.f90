!> ISO_C_BINDING interfaces to the engine's C ABI (include/plbm.h).
!> Buffer ids are 0-based at this level; the lattice_grid shim maps its 1-based
!> iold/inew/imid onto them.
module plbm_capi

   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: PLBM_OK = 0
   integer(c_int), parameter, public :: PLBM_BGK = 0, PLBM_TRT = 1, PLBM_RR = 2, &
                                        PLBM_BGK_SPLIT = 3
   integer(c_int), parameter, public :: PLBM_F64 = 0, PLBM_F64S = 1, PLBM_F32S = 2, PLBM_F16S = 3
   integer(c_int), parameter, public :: PLBM_NRED = 8

   public :: plbm_create, plbm_destroy, plbm_set_params, plbm_upload_f, plbm_download_f
   public :: plbm_equilibrium_init, plbm_step, plbm_run, plbm_stream, plbm_collide, plbm_copy
   public :: plbm_stream_op, plbm_step_op, plbm_set_steps_per_launch, plbm_run_op
   public :: plbm_dugks_collide, plbm_dugks_stream, plbm_dugks_step
   public :: plbm_dugks_stream_faces, plbm_dugks_step_faces
   integer(c_int), parameter, public :: PLBM_STREAM_EXACT = 0, PLBM_STREAM_FVM_BARDOW = 1, &
                                        PLBM_STREAM_FDM_BARDOW = 2, PLBM_STREAM_FDM_SOFONEA = 3
   ! stream_fdm_bardow's stencils selected at compile time (src/fvm_bardow.F90:591-669)
   integer(c_int), parameter, public :: PLBM_STREAM_FDM_BARDOW_WLS = 4, &
                                        PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1 = 5, &
                                        PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2 = 6, &
                                        PLBM_STREAM_FDM_BARDOW_ISO = 7
   public :: plbm_macros, plbm_reduce, plbm_sync, plbm_last_error, plbm_check
   public :: plbm_lambda_d, plbm_magic_number

   interface
      integer(c_int) function plbm_create(nx, ny, nf, storage, out) bind(c, name="plbm_create")
         import :: c_int, c_ptr
         integer(c_int), value :: nx, ny, nf, storage
         type(c_ptr), intent(out) :: out
      end function

      integer(c_int) function plbm_destroy(h) bind(c, name="plbm_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function

      integer(c_int) function plbm_set_params(h, omega, trt_magic) bind(c, name="plbm_set_params")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), value :: omega, trt_magic
      end function

      integer(c_int) function plbm_upload_f(h, buf, f, ld) bind(c, name="plbm_upload_f")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: buf, ld
         real(c_double), intent(in) :: f(*)
      end function

      integer(c_int) function plbm_download_f(h, buf, f, ld) bind(c, name="plbm_download_f")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: buf, ld
         real(c_double), intent(inout) :: f(*)
      end function

      integer(c_int) function plbm_equilibrium_init(h, buf, rho, ux, uy) &
            bind(c, name="plbm_equilibrium_init")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: buf
         real(c_double), intent(in) :: rho(*), ux(*), uy(*)
      end function

      integer(c_int) function plbm_step(h, model, src, dst) bind(c, name="plbm_step")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: model, src, dst
      end function

      integer(c_int) function plbm_run(h, model, nsteps, iold, inew) bind(c, name="plbm_run")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: model, nsteps
         integer(c_int), intent(inout) :: iold, inew
      end function

      integer(c_int) function plbm_stream(h, src, dst) bind(c, name="plbm_stream")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: src, dst
      end function

      integer(c_int) function plbm_stream_op(h, op, src, dst, dt) bind(c, name="plbm_stream_op")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: op, src, dst
         real(c_double), value :: dt
      end function

      integer(c_int) function plbm_set_steps_per_launch(h, model, steps) &
            bind(c, name="plbm_set_steps_per_launch")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: model, steps
      end function

      integer(c_int) function plbm_run_op(h, op, model, nsteps, iold, inew, dt) &
            bind(c, name="plbm_run_op")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: op, model, nsteps
         integer(c_int), intent(inout) :: iold, inew
         real(c_double), value :: dt
      end function

      integer(c_int) function plbm_step_op(h, op, model, src, dst, dt) bind(c, name="plbm_step_op")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: op, model, src, dst
         real(c_double), value :: dt
      end function

      integer(c_int) function plbm_dugks_collide(h, src, dst, omega_full, omega_half) &
            bind(c, name="plbm_dugks_collide")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: src, dst
         real(c_double), value :: omega_full, omega_half
      end function

      integer(c_int) function plbm_dugks_stream(h, src, dst, dt) bind(c, name="plbm_dugks_stream")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: src, dst
         real(c_double), value :: dt
      end function

      integer(c_int) function plbm_dugks_step(h, iold, inew, nsteps, dt, omega_full, &
            omega_half) bind(c, name="plbm_dugks_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: iold, inew, nsteps
         real(c_double), value :: dt, omega_full, omega_half
      end function

      integer(c_int) function plbm_dugks_stream_faces(h, src, dst, dt, omega_face) &
            bind(c, name="plbm_dugks_stream_faces")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: src, dst
         real(c_double), value :: dt, omega_face
      end function

      integer(c_int) function plbm_dugks_step_faces(h, iold, inew, nsteps, dt, omega_full, &
            omega_half, omega_face) bind(c, name="plbm_dugks_step_faces")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: iold, inew, nsteps
         real(c_double), value :: dt, omega_full, omega_half, omega_face
      end function

      integer(c_int) function plbm_collide(h, model, buf) bind(c, name="plbm_collide")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: model, buf
      end function

      integer(c_int) function plbm_copy(h, src, dst) bind(c, name="plbm_copy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: src, dst
      end function

      integer(c_int) function plbm_macros(h, buf, rho, ux, uy) bind(c, name="plbm_macros")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: buf
         real(c_double), intent(out) :: rho(*), ux(*), uy(*)
      end function

      integer(c_int) function plbm_reduce(h, buf, uxa, uya, out) bind(c, name="plbm_reduce")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: buf
         type(c_ptr), value :: uxa, uya
         real(c_double), intent(out) :: out(*)
      end function

      integer(c_int) function plbm_sync(h) bind(c, name="plbm_sync")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function

      type(c_ptr) function plbm_last_error_c() bind(c, name="plbm_last_error")
         import :: c_ptr
      end function

      real(c_double) function plbm_lambda_d(omega, x) bind(c, name="plbm_lambda_d")
         import :: c_double
         real(c_double), value :: omega, x
      end function

      real(c_double) function plbm_magic_number(le, ld) bind(c, name="plbm_magic_number")
         import :: c_double
         real(c_double), value :: le, ld
      end function
   end interface

contains

   !> Text of the last engine error on this thread.
   function plbm_last_error() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n
      p = plbm_last_error_c()
      if (.not. c_associated(p)) then
         msg = ""
         return
      end if
      call c_f_pointer(p, chars, [4096])
      n = 0
      do while (n < 4096)
         if (chars(n+1) == c_null_char) exit
         n = n + 1
      end do
      allocate(character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function

   !> Map a non-zero status to `error stop` with the engine's message (the reference's
   !> own fatal-error convention, src/fvm_bardow.F90:913-916).
   subroutine plbm_check(stat, where)
      integer(c_int), intent(in) :: stat
      character(len=*), intent(in) :: where
      if (stat /= PLBM_OK) then
         write(*,'(A)') "[" // where // "] plbm error: " // plbm_last_error()
         error stop
      end if
   end subroutine

end module
