! DUGKS through the reference API (src/periodic_dugks.F90): the fused step
! (perform_dugks_step with dugks_collide + dugks_stream in the slots), the operators called
! one by one through the slots, and a change of dt.  Built twice: against the reference
! modules (oracle/Makefile target dugks_check_ref, run when the golden fixtures are made)
! and against the engine's Fortran modules (build.py); the raw macroscopic fields it
! writes must be byte-identical.
program dugks_check
   use fvm_bardow
   use periodic_dugks
   implicit none
   type(lattice_grid) :: grid
   integer :: x, y, step, itmp
   real(wp) :: xx, yy

   call alloc_grid(grid, 48, 32, nf=2, log=.false.)
   call grid%set_output_folder(foldername="out")
   call set_properties(grid, nu=0.01_wp, dt=0.5_wp)
   grid%collision => dugks_collide
   grid%streaming => dugks_stream
   do x = 1, grid%nx
      xx = (x - 1) + 0.5_wp
      do y = 1, grid%ny
         yy = (y - 1) + 0.5_wp
         grid%rho(y,x) = 1.0_wp + 1.0e-3_wp*sin(0.3_wp*xx)*cos(0.2_wp*yy)
         grid%ux(y,x) = 0.05_wp*cos(0.25_wp*xx + 0.1_wp*yy)
         grid%uy(y,x) = -0.03_wp*sin(0.15_wp*xx - 0.35_wp*yy)
      end do
   end do
   call set_pdf_to_equilibrium(grid)

   do step = 1, 41
      call perform_dugks_step(grid)
   end do
   call dump("out/lag.bin")          ! update_macros reads f(inew): \bar f^+ of the last step
   itmp = grid%inew
   grid%inew = grid%iold
   call dump("out/latest.bin")       ! f(iold): the latest populations
   grid%inew = itmp

   do step = 1, 3
      call grid%collision()
      call grid%streaming()
      itmp = grid%iold
      grid%iold = grid%inew
      grid%inew = itmp
   end do
   call dump("out/slots.bin")

   call set_properties(grid, nu=0.02_wp, dt=1.0_wp)
   do step = 1, 6
      call perform_dugks_step(grid)
   end do
   call dump("out/dt1.bin")

   call dealloc_grid(grid)
   print *, "dugks fields written"

contains

   subroutine dump(name)
      character(len=*), intent(in) :: name
      integer :: u
      call update_macros(grid)
      open(newunit=u, file=name, access="stream", form="unformatted", status="replace")
      write(u) grid%rho, grid%ux, grid%uy
      close(u)
   end subroutine

end program
