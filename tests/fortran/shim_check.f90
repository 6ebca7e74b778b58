! Semantics of the engine's Fortran modules beyond the unchanged app: the dispatch of
! user procedures in the grid%collision slot (host mirror), perform_triple_step, the
! reduced-precision storage argument and the FP64 reductions.  Prints one line per check
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
   type(lattice_grid) :: a, b, c, d
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
   ! and its third buffer holds the pre-collision (streamed) state of the last step
   err = abs(sum(c%f(1:n,:,:,c%imid)) - sum(c%f(1:n,:,:,c%iold)))
   call check("perform_triple_step keeps the pre-collision mass", err < 1.0e-10_wp, err)

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
