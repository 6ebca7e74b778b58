!> Lattice container of the GPU engine, under the reference's module name and API
!> (ivan-pi/periodic-lbm src/fvm_bardow.F90:13-33), so programs written against the
!> reference compile unchanged.
!>
!> Populations live on the GPU (libplbm.so, include/plbm.h).  `grid%f` keeps the
!> reference's shape f(ld,nx,0:8,nf) as a host mirror; which side holds the current
!> state is tracked per grid:
!>   * the engine's operators (lbm_stream, collide_bgk/trt/rr and the fused step that
!>     perform_lbm_step / perform_step launch for that pair) run on the device and
!>     first upload the host mirror if the host side is the current one;
!>   * any other procedure assigned to grid%streaming / grid%collision works on the host
!>     mirror, so the dispatchers download the device state before calling it;
!>   * update_macros computes rho/ux/uy on the device and copies them into grid%mf,
!>     exactly as the reference fills grid%rho/ux/uy;
!>   * code that reads or writes grid%f directly calls plbm_sync_host(grid) before
!>     reading and plbm_sync_device(grid) after writing.
module fvm_bardow

   use iso_c_binding
   use precision
   use plbm_capi

   implicit none
   private

   public :: wp
   public :: lattice_grid

   public :: alloc_grid, dealloc_grid
   public :: set_properties
   public :: perform_step, perform_triple_step
   public :: update_macros
   public :: output_gnuplot, output_vtk, output_npy
   public :: set_pdf_to_equilibrium
   public :: equilibrium
   public :: stream_fdm_bardow, stream_fvm_bardow, stream_fdm_sofonea
   public :: cx, cy, csqr

   ! engine extensions (prefixed so they cannot clash with program names)
   public :: plbm_sync_host, plbm_sync_device, plbm_fused_model, plbm_perform_steps
   public :: plbm_set_steps_per_pass, plbm_perform_dugks_steps
   public :: plbm_lbm_stream, plbm_collide_bgk, plbm_collide_trt, plbm_collide_rr
   public :: plbm_collide_bgk_split
   public :: plbm_reductions
   public :: plbm_grid_dugks_collide, plbm_grid_dugks_stream, plbm_grid_perform_dugks_step
   public :: plbm_tau_over_dt
   public :: PLBM_F64, PLBM_F64S, PLBM_F32S, PLBM_F16S, PLBM_BGK_SPLIT
   public :: PLBM_BGK, PLBM_TRT, PLBM_RR

   integer, parameter :: WHERE_HOST = 0, WHERE_DEVICE = 1

   ! DUGKS with face collisions, chosen when this module is compiled with the reference's
   ! -DDUGKS (src/periodic_dugks.F90:68,199,267,308): the half collision at 0.75 of its
   ! rate, and the stream relaxes the face values at rate 1/(4 tau/dt + 1)
#if defined(DUGKS)
   logical, parameter :: DUGKS_FACES = .true.
#else
   logical, parameter :: DUGKS_FACES = .false.
#endif

   ! stream_fdm_bardow's stencil, chosen when this module is compiled, with the reference's
   ! preprocessor options (src/fvm_bardow.F90:591-669)
#if defined(FDM_WLS)
   integer(c_int), parameter :: FDM_BARDOW_OP = PLBM_STREAM_FDM_BARDOW_WLS
#elif defined(FDM_WLS_GAUSS_V1)
   integer(c_int), parameter :: FDM_BARDOW_OP = PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1
#elif defined(FDM_WLS_GAUSS_V2)
   integer(c_int), parameter :: FDM_BARDOW_OP = PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2
#elif defined(FDM_ISO)
   integer(c_int), parameter :: FDM_BARDOW_OP = PLBM_STREAM_FDM_BARDOW_ISO
#else
   integer(c_int), parameter :: FDM_BARDOW_OP = PLBM_STREAM_FDM_BARDOW
#endif

   type :: lattice_grid

      integer :: nx, ny

      real(wp), allocatable :: f(:,:,:,:)

      !> Macroscopic fields (contiguous storage) for simple output
      real(wp), allocatable :: mf(:,:,:)

      !> Views of the macroscopic fields
      real(wp), pointer :: rho(:,:) => null()
      real(wp), pointer ::  ux(:,:) => null()
      real(wp), pointer ::  uy(:,:) => null()

      real(wp) :: nu, dt, tau
      real(wp) :: omega, trt_magic
      real(wp) :: csqr

      integer :: iold, inew, imid

      procedure(collision_interface), pointer, pass(grid) :: collision => null()
      procedure(streaming_interface), pointer, pass(grid) :: streaming => null()

      character(len=:), allocatable :: filename, foldername

      character(len=:), allocatable :: logfile
      procedure(gridlog_interface), pointer, pass(grid) :: logger => null()
      integer :: logunit

      ! ---- engine state (not in the reference type; invisible to its users)
      type(c_ptr) :: dev = c_null_ptr     !< plbm_lattice handle
      integer :: storage = PLBM_F64       !< population storage on the device
      integer :: current = WHERE_DEVICE   !< side holding the current populations
      logical :: host_valid = .false.     !< host mirror equals the device state
   contains
      procedure :: set_output_folder
   end type

   abstract interface
      subroutine collision_interface(grid)
         import lattice_grid
         class(lattice_grid), intent(inout) :: grid
      end subroutine
      subroutine streaming_interface(grid)
         import lattice_grid
         class(lattice_grid), intent(inout) :: grid
      end subroutine
      subroutine gridlog_interface(grid, step)
         import lattice_grid
         class(lattice_grid), intent(in) :: grid
         integer, intent(in) :: step
      end subroutine
   end interface

   ! D2Q9 velocities, weights, speed of sound (src/fvm_bardow.F90:87-95)
   real(wp), parameter :: cx(0:8) = [real(wp) :: 0, 1, 0, -1, 0, 1, -1, -1, 1]
   real(wp), parameter :: cy(0:8) = [real(wp) :: 0, 0, 1, 0, -1, 1, 1, -1, -1]

   real(wp), parameter :: w0 = 4._wp / 9._wp, &
                          ws = 1._wp / 9._wp, &
                          wd = 1._wp / 36._wp

   real(wp), parameter :: csqr = 1._wp/3._wp
   real(wp), parameter :: invcsqr = 1._wp/csqr

   character(*), parameter :: FMT_REAL = '(ES24.16E3)'
   character(*), parameter :: FMT_VECTOR = '(3'//FMT_REAL//')'

contains

   !> Second-order equilibrium (src/fvm_bardow.F90:99-126), host side.  The device
   !> kernels evaluate the same expression in the same order.
   pure function equilibrium(rho,ux,uy) result(feq)
      real(wp), intent(in) :: rho, ux, uy
      real(wp) :: feq(0:8)
      real(wp) :: uxx, uyy, indp, uxpy, uxmy

      uxx = ux*ux
      uyy = uy*uy
      indp = 1.0_wp - 1.5_wp*(uxx + uyy)

      feq(0) = w0*rho*indp
      feq(1) = ws*rho*(indp + 3.0_wp*ux + 4.5_wp*uxx)
      feq(2) = ws*rho*(indp + 3.0_wp*uy + 4.5_wp*uyy)
      feq(3) = ws*rho*(indp - 3.0_wp*ux + 4.5_wp*uxx)
      feq(4) = ws*rho*(indp - 3.0_wp*uy + 4.5_wp*uyy)
      uxpy = ux + uy
      feq(5) = wd*rho*(indp + 3.0_wp*uxpy + 4.5_wp*uxpy*uxpy)
      feq(7) = wd*rho*(indp - 3.0_wp*uxpy + 4.5_wp*uxpy*uxpy)
      uxmy = ux - uy
      feq(6) = wd*rho*(indp - 3.0_wp*uxmy + 4.5_wp*uxmy*uxmy)
      feq(8) = wd*rho*(indp + 3.0_wp*uxmy + 4.5_wp*uxmy*uxmy)
   end function

   !> alloc_grid (src/fvm_bardow.F90:129-202) plus the device lattice.
   !> `storage` (engine extension) selects PLBM_F64 (default, reference arithmetic) or
   !> the shifted PLBM_F64S / PLBM_F32S / PLBM_F16S population storage.
   subroutine alloc_grid(grid,nx,ny,nf,log,storage)
      type(lattice_grid), intent(out), target :: grid
      integer, intent(in) :: nx, ny
      integer, intent(in), optional :: nf
      logical, intent(in), optional :: log
      integer, intent(in), optional :: storage

      integer :: nf_, ld
      logical :: log_
      character(len=:), allocatable :: logfile_

      grid%nx = nx
      grid%ny = ny

      ! leading dimension rounded up to a multiple of 16, as in the reference
      ld = ny
      if (mod(ld, 16) > 0) ld = ld + (16 - mod(ld, 16))

      nf_ = 2
      if (present(nf)) nf_ = nf

      ! host mirror: allocated but only touched when synchronised
      allocate(grid%f(ld,nx,0:8,nf_))
      allocate(grid%mf(ny,nx,3))

      grid%rho => grid%mf(:,:,1)
      grid%ux  => grid%mf(:,:,2)
      grid%uy  => grid%mf(:,:,3)

      grid%inew = 1
      grid%iold = 2
      if (nf_ > 2) then
         grid%imid = 3
      else
         grid%imid = -1
      end if

      grid%storage = PLBM_F64
      if (present(storage)) grid%storage = storage
      call plbm_check(plbm_create(int(nx,c_int), int(ny,c_int), int(nf_,c_int), &
         int(grid%storage,c_int), grid%dev), "alloc_grid")
      grid%current = WHERE_DEVICE
      grid%host_valid = .false.

      log_ = .true.
      if (present(log)) log_ = log
      if (log_) then
         logfile_ = "lattice_grid_log.txt"
         if (allocated(grid%logfile)) logfile_ = grid%logfile
         open(newunit=grid%logunit,file=logfile_,status='unknown')
      end if
   end subroutine

   !> dealloc_grid (src/fvm_bardow.F90:204-240)
   subroutine dealloc_grid(grid)
      type(lattice_grid), intent(inout) :: grid
      logical :: isopen

      inquire(grid%logunit,opened=isopen)
      if (isopen) close(grid%logunit)

      if (c_associated(grid%dev)) then
         call plbm_check(plbm_destroy(grid%dev), "dealloc_grid")
         grid%dev = c_null_ptr
      end if

      nullify(grid%uy)
      nullify(grid%ux)
      nullify(grid%rho)
      call deallocate_all_(grid)

   contains

      subroutine deallocate_all_(grid)
         type(lattice_grid), intent(out) :: grid
         associate(nx => grid%nx)
            return
         end associate
      end subroutine

   end subroutine

   !> set_properties (src/fvm_bardow.F90:242-269): tau = nu/csqr, omega = dt/(tau+dt/2),
   !> trt_magic = magic or (tau/dt)**2.
   subroutine set_properties(grid, nu, dt, magic)
      type(lattice_grid), intent(inout) :: grid
      real(wp), intent(in) :: nu, dt
      real(wp), optional :: magic

      real(wp) :: tau

      grid%nu = nu
      grid%dt = dt
      grid%csqr = csqr
      tau = invcsqr*nu
      grid%tau = tau
      grid%omega = dt/(tau + 0.5_wp*dt)

      if (present(magic)) then
         grid%trt_magic = magic
      else
         grid%trt_magic = (tau/dt)**2
      end if

      print *, "trt magic = ", grid%trt_magic

      if (c_associated(grid%dev)) then
         call plbm_check(plbm_set_params(grid%dev, grid%omega, grid%trt_magic), "set_properties")
      end if
   end subroutine

   ! ------------------------------------------------------------------ synchronisation

   !> Make the host mirror grid%f equal to the current populations.
   subroutine plbm_sync_host(grid)
      class(lattice_grid), intent(inout) :: grid
      integer :: b
      if (grid%current == WHERE_HOST .or. grid%host_valid) return
      call plbm_check(plbm_sync(grid%dev), "plbm_sync_host")
      do b = 1, size(grid%f, 4)
         call plbm_check(plbm_download_f(grid%dev, int(b-1,c_int), grid%f(:,:,:,b), &
            int(size(grid%f,1),c_int)), "plbm_sync_host")
      end do
      grid%host_valid = .true.
   end subroutine

   !> Upload grid%f after the host mirror was written directly.
   subroutine plbm_sync_device(grid)
      class(lattice_grid), intent(inout) :: grid
      integer :: b
      do b = 1, size(grid%f, 4)
         call plbm_check(plbm_upload_f(grid%dev, int(b-1,c_int), grid%f(:,:,:,b), &
            int(size(grid%f,1),c_int)), "plbm_sync_device")
      end do
      grid%current = WHERE_DEVICE
      grid%host_valid = .true.
   end subroutine

   subroutine require_device(grid)
      class(lattice_grid), intent(inout) :: grid
      if (grid%current == WHERE_HOST) call plbm_sync_device(grid)
      call plbm_check(plbm_set_params(grid%dev, grid%omega, grid%trt_magic), "plbm")
   end subroutine

   subroutine device_written(grid)
      class(lattice_grid), intent(inout) :: grid
      grid%current = WHERE_DEVICE
      grid%host_valid = .false.
   end subroutine

   subroutine require_host(grid)
      class(lattice_grid), intent(inout) :: grid
      call plbm_sync_host(grid)
      grid%current = WHERE_HOST
   end subroutine

   ! ------------------------------------------------------------------ device operators

   !> lbm_stream (src/periodic_lbm.f90:32-129) on the device: f(iold) -> f(inew).
   subroutine plbm_lbm_stream(grid)
      class(lattice_grid), intent(inout) :: grid
      call require_device(grid)
      call plbm_check(plbm_stream(grid%dev, int(grid%iold-1,c_int), int(grid%inew-1,c_int)), &
         "lbm_stream")
      call device_written(grid)
   end subroutine

   subroutine collide_on_device(grid, model, name)
      class(lattice_grid), intent(inout) :: grid
      integer(c_int), intent(in) :: model
      character(len=*), intent(in) :: name
      call require_device(grid)
      call plbm_check(plbm_collide(grid%dev, model, int(grid%inew-1,c_int)), name)
      call device_written(grid)
   end subroutine

   !> collide_bgk (src/collision_bgk.F90:17-82) on the device, in place on f(inew).
   subroutine plbm_collide_bgk(grid)
      class(lattice_grid), intent(inout) :: grid
      call collide_on_device(grid, PLBM_BGK, "collide_bgk")
   end subroutine

   !> collide_bgk of the -DSPLIT build (bgk_kernel_cache, src/collision_bgk.F90:84-176)
   !> on the device, in place on f(inew).
   subroutine plbm_collide_bgk_split(grid)
      class(lattice_grid), intent(inout) :: grid
      call collide_on_device(grid, PLBM_BGK_SPLIT, "collide_bgk")
   end subroutine

   !> collide_trt (src/collision_trt.F90:41-160) on the device, in place on f(inew).
   subroutine plbm_collide_trt(grid)
      class(lattice_grid), intent(inout) :: grid
      call collide_on_device(grid, PLBM_TRT, "collide_trt")
   end subroutine

   !> collide_rr (src/collision_regularized.F90:21-202) on the device, in place on f(inew).
   subroutine plbm_collide_rr(grid)
      class(lattice_grid), intent(inout) :: grid
      call collide_on_device(grid, PLBM_RR, "collide_rr")
   end subroutine

   !> Model id of the fused kernel for the grid's (streaming, collision) pair, or -1 when
   !> the pair is not an engine streaming operator (lbm_stream or a FV/FD operator) with
   !> collide_bgk/trt/rr (then the step is dispatched through the procedure pointers).
   integer function plbm_fused_model(grid) result(model)
      type(lattice_grid), intent(in) :: grid
      model = -1
      if (plbm_fused_op(grid) < 0) return
      if (associated(grid%collision, plbm_collide_bgk)) then
         model = PLBM_BGK
      else if (associated(grid%collision, plbm_collide_trt)) then
         model = PLBM_TRT
      else if (associated(grid%collision, plbm_collide_rr)) then
         model = PLBM_RR
      else if (associated(grid%collision, plbm_collide_bgk_split)) then
         model = PLBM_BGK_SPLIT
      end if
   end function

   !> Streaming-operator id (PLBM_STREAM_*) of grid%streaming, or -1 if it is not one of
   !> the engine's streaming operators.
   integer function plbm_fused_op(grid) result(op)
      type(lattice_grid), intent(in) :: grid
      op = -1
      if (associated(grid%streaming, plbm_lbm_stream)) then
         op = PLBM_STREAM_EXACT
      else if (associated(grid%streaming, stream_fvm_bardow)) then
         op = PLBM_STREAM_FVM_BARDOW
      else if (associated(grid%streaming, stream_fdm_bardow)) then
         op = FDM_BARDOW_OP
      else if (associated(grid%streaming, stream_fdm_sofonea)) then
         op = PLBM_STREAM_FDM_SOFONEA
      end if
   end function

   logical function is_engine_op_s(grid) result(yes)
      type(lattice_grid), intent(in) :: grid
      yes = associated(grid%streaming, plbm_lbm_stream) .or. &
            associated(grid%streaming, stream_fvm_bardow) .or. &
            associated(grid%streaming, stream_fdm_bardow) .or. &
            associated(grid%streaming, stream_fdm_sofonea) .or. &
            associated(grid%streaming, plbm_grid_dugks_stream)
   end function

   logical function is_engine_op_c(grid) result(yes)
      type(lattice_grid), intent(in) :: grid
      yes = associated(grid%collision, plbm_collide_bgk) .or. &
            associated(grid%collision, plbm_collide_trt) .or. &
            associated(grid%collision, plbm_collide_rr) .or. &
            associated(grid%collision, plbm_collide_bgk_split) .or. &
            associated(grid%collision, plbm_grid_dugks_collide)
   end function

   !> Call grid%streaming, moving the populations to the side the operator works on.
   subroutine dispatch_streaming(grid)
      type(lattice_grid), intent(inout) :: grid
      if (.not. is_engine_op_s(grid)) call require_host(grid)
      call grid%streaming()
   end subroutine

   subroutine dispatch_collision(grid)
      type(lattice_grid), intent(inout) :: grid
      if (.not. is_engine_op_c(grid)) call require_host(grid)
      call grid%collision()
   end subroutine

   !> One stream + collide step with the reference's buffer swap
   !> (src/periodic_lbm.f90:15-29, src/fvm_bardow.F90:307-320).  A pair of engine
   !> operators runs as one fused kernel; other pairs go through the procedure pointers.
   subroutine perform_step(grid)
      type(lattice_grid), intent(inout) :: grid
      integer :: model, itmp

      model = plbm_fused_model(grid)
      if (model >= 0) then
         call require_device(grid)
         call plbm_check(plbm_step_op(grid%dev, int(plbm_fused_op(grid),c_int), &
            int(model,c_int), int(grid%iold-1,c_int), int(grid%inew-1,c_int), grid%dt), &
            "perform_step")
         call device_written(grid)
      else
         call dispatch_streaming(grid)
         call dispatch_collision(grid)
      end if

      itmp = grid%iold
      grid%iold = grid%inew
      grid%inew = itmp
   end subroutine

   !> Engine extension: steps per pass of plbm_perform_steps for one collision model
   !> (PLBM_BGK, PLBM_TRT, PLBM_RR, PLBM_BGK_SPLIT); 0 = chosen from the grid size.
   subroutine plbm_set_steps_per_pass(grid, model, steps)
      type(lattice_grid), intent(inout) :: grid
      integer, intent(in) :: model, steps
      call plbm_check(plbm_set_steps_per_launch(grid%dev, int(model,c_int), int(steps,c_int)), &
         "plbm_set_steps_per_pass")
   end subroutine

   !> Engine extension: the effect of `nsteps` calls of perform_step in one call.  With an
   !> engine streaming operator (lbm_stream, stream_fvm_bardow, stream_fdm_bardow,
   !> stream_fdm_sofonea) and an engine collision in the slots the steps run in temporally
   !> blocked passes of several steps per sweep over device memory (plbm_run_op;
   !> bit-identical to single steps in FP64 storage, the slots end as after nsteps
   !> perform_step calls); any other pair is stepped one perform_step at a time.
   subroutine plbm_perform_steps(grid, nsteps)
      type(lattice_grid), intent(inout) :: grid
      integer, intent(in) :: nsteps
      integer :: model, s
      integer(c_int) :: iold, inew

      model = plbm_fused_model(grid)
      if (model >= 0) then
         call require_device(grid)
         iold = int(grid%iold - 1, c_int)
         inew = int(grid%inew - 1, c_int)
         call plbm_check(plbm_run_op(grid%dev, int(plbm_fused_op(grid),c_int), &
            int(model,c_int), int(nsteps,c_int), iold, inew, grid%dt), "plbm_perform_steps")
         call device_written(grid)
         grid%iold = iold + 1
         grid%inew = inew + 1
      else
         do s = 1, nsteps
            call perform_step(grid)
         end do
      end if
   end subroutine

   !> perform_triple_step (src/fvm_bardow.F90:322-340): keeps the pre-collision state.
   subroutine perform_triple_step(grid)
      type(lattice_grid), intent(inout) :: grid
      integer :: itmp

      call dispatch_streaming(grid)
      if (grid%current == WHERE_DEVICE) then
         call plbm_check(plbm_copy(grid%dev, int(grid%inew-1,c_int), int(grid%iold-1,c_int)), &
            "perform_triple_step")
         call device_written(grid)
      else
         grid%f(:,:,:,grid%iold) = grid%f(:,:,:,grid%inew)
      end if
      call dispatch_collision(grid)

      itmp = grid%iold
      grid%iold = grid%inew
      grid%inew = grid%imid
      grid%imid = itmp
   end subroutine

   !> set_pdf_to_equilibrium (src/fvm_bardow.F90:272-305): f(iold) = feq(rho, u).
   subroutine set_pdf_to_equilibrium(grid)
      type(lattice_grid), intent(inout) :: grid
      if (grid%current == WHERE_HOST) call plbm_sync_device(grid)
      call plbm_check(plbm_equilibrium_init(grid%dev, int(grid%iold-1,c_int), &
         grid%rho, grid%ux, grid%uy), "set_pdf_to_equilibrium")
      call device_written(grid)
   end subroutine

   !> update_macros (src/fvm_bardow.F90:343-390): rho, u from f(inew) into grid%mf.
   !> (After perform_lbm_step, inew holds the state one step before the latest.)
   subroutine update_macros(grid)
      type(lattice_grid), intent(inout) :: grid
      integer :: x, y
      real(wp) :: fs(0:8), rho, invrho

      if (grid%current == WHERE_DEVICE) then
         call plbm_check(plbm_macros(grid%dev, int(grid%inew-1,c_int), &
            grid%rho, grid%ux, grid%uy), "update_macros")
      else
         do x = 1, grid%nx
            do y = 1, grid%ny
               fs = grid%f(y,x,:,grid%inew)
               rho = fs(0) + (((fs(5) + fs(7)) + (fs(6) + fs(8))) + &
                      ((fs(1) + fs(3)) + (fs(2) + fs(4))))
               grid%rho(y,x) = rho
               invrho = 1.0_wp/rho
               grid%ux(y,x) = invrho * (((fs(5) - fs(7)) + (fs(8) - fs(6))) + (fs(1) - fs(3)))
               grid%uy(y,x) = invrho * (((fs(5) - fs(7)) + (fs(6) - fs(8))) + (fs(2) - fs(4)))
            end do
         end do
      end if
   end subroutine

   !> FP64 device reductions of f(buf) (engine extension): out = [mass, kinetic energy,
   !> sum|u-ua|^2, sum|ua|^2, max|u|, non-finite cells, sum|u|^2, 0].
   subroutine plbm_reductions(grid, buf, out, uxa, uya)
      type(lattice_grid), intent(inout) :: grid
      integer, intent(in) :: buf
      real(wp), intent(out) :: out(0:7)
      real(wp), intent(in), optional, target, contiguous :: uxa(:,:), uya(:,:)
      if (grid%current == WHERE_HOST) call plbm_sync_device(grid)
      if (present(uxa) .and. present(uya)) then
         call plbm_check(plbm_reduce(grid%dev, int(buf-1,c_int), c_loc(uxa), c_loc(uya), out), &
            "plbm_reductions")
      else
         call plbm_check(plbm_reduce(grid%dev, int(buf-1,c_int), c_null_ptr, c_null_ptr, out), &
            "plbm_reductions")
      end if
   end subroutine

   ! ------------------------------------------------------------------ non-LBM streaming

   subroutine stream_op_on_device(grid, op, name)
      class(lattice_grid), intent(inout) :: grid
      integer(c_int), intent(in) :: op
      character(len=*), intent(in) :: name
      call require_device(grid)
      call plbm_check(plbm_stream_op(grid%dev, op, int(grid%iold-1,c_int), &
         int(grid%inew-1,c_int), grid%dt), name)
      call device_written(grid)
   end subroutine

   !> stream_fvm_bardow (src/fvm_bardow.F90:393-509): finite-volume streaming with
   !> bilinear face interpolation, f(iold) -> f(inew), time step grid%dt, on the device.
   subroutine stream_fvm_bardow(grid)
      class(lattice_grid), intent(inout) :: grid
      call stream_op_on_device(grid, PLBM_STREAM_FVM_BARDOW, "stream_fvm_bardow")
   end subroutine

   !> stream_fdm_bardow (src/fvm_bardow.F90:511-685): Lax-Wendroff finite-difference
   !> streaming on the device, with the stencil this module was compiled for (central
   !> differences, or -DFDM_WLS / -DFDM_WLS_GAUSS_V1 / -DFDM_WLS_GAUSS_V2 / -DFDM_ISO).
   subroutine stream_fdm_bardow(grid)
      class(lattice_grid), intent(inout) :: grid
      call stream_op_on_device(grid, FDM_BARDOW_OP, "stream_fdm_bardow")
   end subroutine

   !> stream_fdm_sofonea (src/fvm_bardow.F90:688-893): 1-D Lax-Wendroff along each
   !> lattice direction on the device.
   subroutine stream_fdm_sofonea(grid)
      class(lattice_grid), intent(inout) :: grid
      call stream_op_on_device(grid, PLBM_STREAM_FDM_SOFONEA, "stream_fdm_sofonea")
   end subroutine

   ! ------------------------------------------------------------------ DUGKS

   !> tau_over_dt (src/periodic_dugks.F90:40-44).
   pure function plbm_tau_over_dt(grid) result(tau_d)
      type(lattice_grid), intent(in) :: grid
      real(wp) :: tau_d
      tau_d = grid%tau/grid%dt
   end function

   !> dugks_collide (src/periodic_dugks.F90:46-77) on the device:
   !> f(inew) = BGK(f(iold), grid%omega), f(iold) = BGK(f(iold), 1/(tau/dt + 1/2)).
   subroutine plbm_grid_dugks_collide(grid)
      class(lattice_grid), intent(inout) :: grid
      call require_device(grid)
      call plbm_check(plbm_dugks_collide(grid%dev, int(grid%iold-1,c_int), &
         int(grid%inew-1,c_int), grid%omega, dugks_omega_half(grid)), "dugks_collide")
      call device_written(grid)
   end subroutine

   !> dugks_stream (src/periodic_dugks.F90:168-299) on the device: f(inew) minus the
   !> face-flux balance of f(iold).
   subroutine plbm_grid_dugks_stream(grid)
      class(lattice_grid), intent(inout) :: grid
      call require_device(grid)
      if (DUGKS_FACES) then
         call plbm_check(plbm_dugks_stream_faces(grid%dev, int(grid%iold-1,c_int), &
            int(grid%inew-1,c_int), grid%dt, dugks_omega_face(grid)), "dugks_stream")
      else
         call plbm_check(plbm_dugks_stream(grid%dev, int(grid%iold-1,c_int), &
            int(grid%inew-1,c_int), grid%dt), "dugks_stream")
      end if
      call device_written(grid)
   end subroutine

   !> Half-step rate of dugks_collide: 1/(tau/dt + 1/2) (src/periodic_dugks.F90:56), times
   !> 0.75 in the -DDUGKS build (:68-71).
   real(wp) function dugks_omega_half(grid) result(omega)
      type(lattice_grid), intent(in) :: grid
      omega = 1.0_wp/(plbm_tau_over_dt(grid) + 0.5_wp)
      if (DUGKS_FACES) omega = 0.75_wp*omega
   end function

   !> Face-collision rate of the -DDUGKS build's dugks_stream (src/periodic_dugks.F90:179).
   real(wp) function dugks_omega_face(grid) result(omega)
      type(lattice_grid), intent(in) :: grid
      omega = 1.0_wp/(4.0_wp*plbm_tau_over_dt(grid) + 1.0_wp)
   end function

   !> The engine's fused DUGKS steps for this build (plbm_dugks_step, or
   !> plbm_dugks_step_faces with -DDUGKS).
   integer(c_int) function dugks_steps(grid, nsteps) result(status)
      type(lattice_grid), intent(in) :: grid
      integer, intent(in) :: nsteps
      if (DUGKS_FACES) then
         status = plbm_dugks_step_faces(grid%dev, int(grid%iold-1,c_int), &
            int(grid%inew-1,c_int), int(nsteps,c_int), grid%dt, grid%omega, &
            dugks_omega_half(grid), dugks_omega_face(grid))
      else
         status = plbm_dugks_step(grid%dev, int(grid%iold-1,c_int), &
            int(grid%inew-1,c_int), int(nsteps,c_int), grid%dt, grid%omega, &
            dugks_omega_half(grid))
      end if
   end function

   !> perform_dugks_step (src/periodic_dugks.F90:25-38): collision, streaming, swap.  With
   !> dugks_collide + dugks_stream in the slots both run as the engine's fused step.
   subroutine plbm_grid_perform_dugks_step(grid)
      type(lattice_grid), intent(inout) :: grid
      integer :: itmp

      if (associated(grid%collision, plbm_grid_dugks_collide) .and. &
          associated(grid%streaming, plbm_grid_dugks_stream)) then
         call require_device(grid)
         call plbm_check(dugks_steps(grid, 1), "perform_dugks_step")
         call device_written(grid)
      else
         call dispatch_collision(grid)
         call dispatch_streaming(grid)
      end if

      itmp = grid%iold
      grid%iold = grid%inew
      grid%inew = itmp
   end subroutine

   !> Engine extension: the effect of `nsteps` perform_dugks_step calls in one call; with
   !> dugks_collide + dugks_stream in the slots all steps but the last run in temporally
   !> blocked passes (bit-identical, same slot contents as the step-by-step calls).
   subroutine plbm_perform_dugks_steps(grid, nsteps)
      type(lattice_grid), intent(inout) :: grid
      integer, intent(in) :: nsteps
      integer :: itmp, s

      if (associated(grid%collision, plbm_grid_dugks_collide) .and. &
          associated(grid%streaming, plbm_grid_dugks_stream)) then
         call require_device(grid)
         call plbm_check(dugks_steps(grid, nsteps), "plbm_perform_dugks_steps")
         call device_written(grid)
         if (mod(nsteps, 2) == 1) then
            itmp = grid%iold
            grid%iold = grid%inew
            grid%inew = itmp
         end if
      else
         do s = 1, nsteps
            call plbm_grid_perform_dugks_step(grid)
         end do
      end if
   end subroutine

   ! ------------------------------------------------------------------ output

   function output_name(grid, step, ext) result(fullname)
      type(lattice_grid), intent(in) :: grid
      integer, intent(in), optional :: step
      character(len=*), intent(in) :: ext
      character(len=:), allocatable :: fullname
      character(len=64) :: istr
      istr = ''
      if (present(step)) write(istr,'(I0.9)') step
      fullname = ''
      if (allocated(grid%foldername)) fullname = trim(grid%foldername) // '/'
      fullname = fullname // grid%filename // trim(istr) // ext
   end function

   subroutine make_folder(grid, who)
      type(lattice_grid), intent(in) :: grid
      character(len=*), intent(in) :: who
      integer :: istat
      if (.not. allocated(grid%foldername)) return
      call execute_command_line("mkdir -p "//trim(grid%foldername), exitstat=istat, wait=.true.)
      if (istat /= 0) then
         write(*,'(A)') "["//who//"] error making directory "//grid%foldername
         error stop
      end if
   end subroutine

   !> output_gnuplot (src/fvm_bardow.F90:895-926, src/output/gnuplot.F90:20-40):
   !> columns x, y, rho, ux, uy per cell, a blank line after each column of cells.
   !> The reference assigns the y coordinate to rx and leaves ry unset
   !> (src/output/gnuplot.F90:31-34); here both coordinates are written.
   subroutine output_gnuplot(grid, step)
      type(lattice_grid), intent(in) :: grid
      integer, intent(in), optional :: step
      integer :: unit, x, y
      real(wp) :: rx, ry

      call make_folder(grid, "output_grid")
      open(newunit=unit, file=output_name(grid, step, '.txt'))
      do x = 1, grid%nx
         rx = (x - 1) + 0.5_wp
         do y = 1, grid%ny
            ry = (y - 1) + 0.5_wp
            write(unit,'(5('//FMT_REAL//',:,1X))') rx, ry, grid%rho(y,x), grid%ux(y,x), grid%uy(y,x)
         end do
         write(unit,*)
      end do
      close(unit)
   end subroutine

   !> output_npy (src/fvm_bardow.F90:930-961, src/output/npy.f90:13-26): grid%mf as an
   !> .npy file (format 1.0, little-endian float64, Fortran order, shape (ny, nx, 3)).
   subroutine output_npy(grid, step)
      type(lattice_grid), intent(in) :: grid
      integer, intent(in), optional :: step
      integer :: unit, hlen, pad
      character(len=:), allocatable :: dict, header
      character(len=32) :: s1, s2

      call make_folder(grid, "output_grid")
      write(s1,'(I0)') grid%ny
      write(s2,'(I0)') grid%nx
      ! header text and padding as written by the reference's npy writer (fortran-lang
      ! stdlib save_npy, format 1.0): every shape entry followed by ", ", then spaces up to
      ! the next 64-byte boundary (a full 64 when already aligned) and a newline
      dict = "{'descr': '<f8', 'fortran_order': True, 'shape': (" // trim(s1) // ", " // &
             trim(s2) // ", 3, ), }"
      hlen = len(dict)
      pad = 64 - mod(hlen + 10 + 1, 64)
      header = dict // repeat(' ', pad) // achar(10)
      open(newunit=unit, file=output_name(grid, step, '.npy'), access='stream', &
         form='unformatted', status='replace')
      write(unit) achar(147) // "NUMPY" // achar(1) // achar(0)
      write(unit) achar(mod(len(header), 256)) // achar(len(header) / 256)
      write(unit) header
      write(unit) grid%mf
      close(unit)
   end subroutine

   !> output_vtk (src/fvm_bardow.F90:963-997, src/output/vtk.F90:154-196): legacy ASCII
   !> STRUCTURED_POINTS with cell data Density and Velocity.
   subroutine output_vtk(grid, step, binary)
      type(lattice_grid), intent(in) :: grid
      integer, intent(in), optional :: step
      logical, intent(in), optional :: binary
      integer :: unit_, i, j
      logical :: binary_

      binary_ = .false.
      if (present(binary)) binary_ = binary
      if (binary_) then
         print *, "binary output not implemented"
         return
      end if

      open(newunit=unit_, file=output_name(grid, step, '.vtk'))
      write(unit_, '(A)') "# vtk DataFile Version 3.0"
      write(unit_, '(A)') "fluid"
      write(unit_, '(A)') "ASCII"
      write(unit_, '(A)') "DATASET STRUCTURED_POINTS"
      write(unit_, '(A,3(I0,1X))') "DIMENSIONS ", grid%nx+1, grid%ny+1, 2
      write(unit_, '(A,'// FMT_VECTOR //')') "ORIGIN  ", 0.0_wp, 0.0_wp, 0.0_wp
      write(unit_, '(A,'// FMT_VECTOR //')') "SPACING ", 1.0_wp, 1.0_wp, 1.0_wp
      write(unit_, *)
      write(unit_, '(A,I0)') "CELL_DATA ", grid%nx*grid%ny
      write(unit_, '(A)') "SCALARS Density float 1"
      write(unit_, '(A)') "LOOKUP_TABLE default"
      do j = 1, grid%ny
         do i = 1, grid%nx
            write(unit_, FMT_REAL) grid%rho(j,i)
         end do
      end do
      write(unit_, *)
      write(unit_, '(A)') "VECTORS Velocity float"
      do j = 1, grid%ny
         do i = 1, grid%nx
            write(unit_, FMT_VECTOR) grid%ux(j,i), grid%uy(j,i), 0.0_wp
         end do
      end do
      close(unit_)
   end subroutine

   !> set_output_folder (src/fvm_bardow.F90:999-1025)
   subroutine set_output_folder(grid,foldername,verbose)
      class(lattice_grid), intent(inout) :: grid
      character(len=*), intent(in) :: foldername
      logical, intent(in), optional :: verbose
      integer :: istat
      character(len=:), allocatable :: mkdir_opts

      mkdir_opts = '-p'
      if (present(verbose)) then
         if (verbose) mkdir_opts = mkdir_opts // 'v'
      end if
      call execute_command_line('mkdir ' // mkdir_opts // ' ' // trim(foldername), &
         exitstat=istat, wait=.true.)
      if (istat /= 0) then
         write(*,'(A)') "[set_output_folder] error making directory "//trim(foldername)
         error stop
      end if
      grid%foldername = foldername
   end subroutine

end module
