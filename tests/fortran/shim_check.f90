! Semantics of the engine's Fortran modules beyond the unchanged app: the dispatch of
! user procedures in the grid%collision slot (host mirror), perform_triple_step, the
! reduced-precision storage argument, the FP64 reductions and plbm_perform_steps.  Prints one line per check
! and stops with an error on the first failure.
module user_ops
   use fvm_bardow
   implicit none
contains
   !> A user-written BGK collision on the host mirror (textbook form, not the engine's)
   subroutine my_bgk(grid)
      class(lattice_grid), intent(inout) :: grid
      integer :: x, y, q
      real(wp) :: f(0:8), rho, ux, uy, feq(0:8)
      do x = 1, grid%nx
         do y = 1, grid%ny
            f = grid%f(y,x,:,grid%inew)
            rho = sum(f)
            ux = sum(f*cx)/rho
            uy = sum(f*cy)/rho
            feq = equilibrium(rho, ux, uy)
            do q = 0, 8
               grid%f(y,x,q,grid%inew) = f(q) + grid%omega*(feq(q) - f(q))
            end do
         end do
      end do
   end subroutine
end module

program shim_check
   use fvm_bardow
   use periodic_lbm, only: perform_lbm_step, lbm_stream
   use collision_bgk, only: collide_bgk
   use collision_trt, only: collide_trt
   use taylor_green, only: taylor_green_t, pi
   use user_ops
   implicit none
   integer, parameter :: n = 48, nsteps = 40
   type(lattice_grid) :: a, b, c, d, e
   type(taylor_green_t) :: tg
   real(wp) :: kx, err, red(0:7), e32
   integer :: s

   kx = 2*pi/n
   tg = taylor_green_t(n, n, kx, kx, 0.05_wp, 0.02_wp)

   call init(a, 2, PLBM_F64)
   call init(b, 2, PLBM_F64)
   call init(c, 3, PLBM_F64)
   call init(d, 2, PLBM_F32S)
   a%collision => collide_bgk
   b%collision => my_bgk
   c%collision => collide_trt
   d%collision => collide_trt

   do s = 1, nsteps
      call perform_lbm_step(a)
      call perform_lbm_step(b)
      call perform_triple_step(c)
      call perform_lbm_step(d)
   end do

   ! 1. user collision on the host mirror == engine collide_bgk (different expression trees)
   call plbm_sync_host(a)
   call plbm_sync_host(b)
   err = maxval(abs(a%f(1:n,:,:,a%iold) - b%f(1:n,:,:,b%iold)))
   call check("user collision through the slot", err < 1.0e-14_wp, err)

   ! 2. perform_triple_step: latest state equals a double-buffer TRT run
   call init(a, 2, PLBM_F64)
   a%collision => collide_trt
   do s = 1, nsteps
      call perform_lbm_step(a)
   end do
   call plbm_sync_host(a)
   call plbm_sync_host(c)
   err = maxval(abs(a%f(1:n,:,:,a%iold) - c%f(1:n,:,:,c%iold)))
   call check("perform_triple_step latest state", err == 0.0_wp, err)
   ! and its third buffer holds the pre-collision (streamed) state of the last step: the
   ! double-buffer run one step short, streamed once more without a collision
   call init(e, 2, PLBM_F64)
   e%collision => collide_trt
   do s = 1, nsteps - 1
      call perform_lbm_step(e)
   end do
   call e%streaming()
   call plbm_sync_host(e)
   err = maxval(abs(e%f(1:n,:,:,e%inew) - c%f(1:n,:,:,c%imid)))
   call check("perform_triple_step third slot = streamed state", err == 0.0_wp, err)
   ! and the slot after it the streamed state of the step before
   call init(e, 2, PLBM_F64)
   e%collision => collide_trt
   do s = 1, nsteps - 2
      call perform_lbm_step(e)
   end do
   call e%streaming()
   call plbm_sync_host(e)
   err = maxval(abs(e%f(1:n,:,:,e%inew) - c%f(1:n,:,:,c%inew)))
   call check("perform_triple_step second slot = previous streamed state", err == 0.0_wp, err)

   ! 3. FP32-shifted storage tracks FP64 (velocities after update_macros)
   call update_macros(a)
   call update_macros(d)
   e32 = maxval(abs(a%ux - d%ux)) / maxval(abs(a%ux))
   call check("FP32-shifted storage vs FP64", e32 < 1.0e-5_wp, e32)

   ! 4. device reductions: mass of the latest state
   call plbm_reductions(a, a%iold, red)
   err = abs(red(0) - sum(a%f(1:n,:,:,a%iold)))/red(0)
   call check("device mass reduction", err < 1.0e-13_wp, err)

   ! 5. a host-side edit of grid%f is picked up after plbm_sync_device
   a%f(3,4,0,a%iold) = a%f(3,4,0,a%iold) + 1.0e-3_wp
   call plbm_sync_device(a)
   call plbm_reductions(a, a%iold, red)
   err = abs(red(0) - sum(a%f(1:n,:,:,a%iold)))/red(0)
   call check("plbm_sync_device after a host edit", err < 1.0e-13_wp, err)

   ! 6. plbm_perform_steps (passes of several steps) == perform_step called one by one,
   !    both buffers, on a grid large enough for passes (240 x 200: ragged strips)
   call check_perform_steps(37)
   call check_perform_steps(40)

   call dealloc_grid(a)
   call dealloc_grid(b)
   call dealloc_grid(c)
   call dealloc_grid(d)
   print *, "shim_check: all checks passed"

contains

   subroutine init(g, nf, storage)
      type(lattice_grid), intent(inout) :: g
      integer, intent(in) :: nf, storage
      if (allocated(g%f)) call dealloc_grid(g)
      call alloc_grid(g, n, n, nf=nf, log=.false., storage=storage)
      call set_properties(g, 0.02_wp, 1.0_wp, 0.25_wp)
      g%streaming => lbm_stream
      call tg%eval(0.0_wp, g%rho, g%ux, g%uy)
      g%rho = g%rho/csqr + 1.0_wp
      call set_pdf_to_equilibrium(g)
   end subroutine

   subroutine check_perform_steps(ns)
      integer, intent(in) :: ns
      type(lattice_grid) :: p, q
      integer :: k, x, y
      character(len=64) :: what
      do k = 1, 2
         if (k == 1) call alloc_grid(p, 240, 200, nf=2, log=.false.)
         if (k == 2) call alloc_grid(q, 240, 200, nf=2, log=.false.)
      end do
      call set_properties(p, 0.01_wp, 1.0_wp, 0.25_wp)
      call set_properties(q, 0.01_wp, 1.0_wp, 0.25_wp)
      call plbm_set_steps_per_pass(p, PLBM_TRT, 4)
      p%streaming => lbm_stream
      q%streaming => lbm_stream
      p%collision => collide_trt
      q%collision => collide_trt
      do x = 1, 240
         do y = 1, 200
            p%rho(y,x) = 1.0_wp + 1.0e-3_wp*sin(0.1_wp*x)*cos(0.2_wp*y)
            p%ux(y,x) = 0.04_wp*cos(0.05_wp*x + 0.1_wp*y)
            p%uy(y,x) = -0.03_wp*sin(0.07_wp*x - 0.03_wp*y)
         end do
      end do
      q%mf = p%mf
      call set_pdf_to_equilibrium(p)
      call set_pdf_to_equilibrium(q)
      call plbm_perform_steps(p, ns)
      do k = 1, ns
         call perform_step(q)
      end do
      call plbm_sync_host(p)
      call plbm_sync_host(q)
      err = max(maxval(abs(p%f(1:200,:,:,p%iold) - q%f(1:200,:,:,q%iold))), &
                maxval(abs(p%f(1:200,:,:,p%inew) - q%f(1:200,:,:,q%inew))))
      write(what, '(A,I0,A)') "plbm_perform_steps(", ns, ") == perform_step x n"
      call check(trim(what), err == 0.0_wp .and. p%iold == q%iold, err)
      call dealloc_grid(p)
      call dealloc_grid(q)
   end subroutine

   subroutine check(what, ok, val)
      character(len=*), intent(in) :: what
      logical, intent(in) :: ok
      real(wp), intent(in) :: val
      if (ok) then
         print '(A,A,ES10.3)', "PASS ", what//": ", val
      else
         print '(A,A,ES10.3)', "FAIL ", what//": ", val
         error stop "shim_check failed"
      end if
   end subroutine
end program
