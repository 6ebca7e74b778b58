! Writes the reference API's output files for a fixed non-square field.  Built twice:
! against the reference modules (oracle/Makefile target outputs_check_ref, run when the
! golden fixtures are made) and against the engine's Fortran modules (build.py); the
! VTK and npy files must be byte-identical.
program outputs_check
   use fvm_bardow
   implicit none
   type(lattice_grid) :: grid
   integer :: x, y
   real(wp) :: xx, yy

   call alloc_grid(grid, 24, 16, nf=2, log=.false.)
   grid%filename = "fields"
   call grid%set_output_folder(foldername="out")
   do x = 1, grid%nx
      xx = (x - 1) + 0.5_wp
      do y = 1, grid%ny
         yy = (y - 1) + 0.5_wp
         grid%rho(y,x) = 1.0_wp + 1.0e-3_wp*sin(0.3_wp*xx)*cos(0.2_wp*yy)
         grid%ux(y,x) = 0.05_wp*cos(0.25_wp*xx + 0.1_wp*yy)
         grid%uy(y,x) = -0.03_wp*sin(0.15_wp*xx - 0.35_wp*yy) + 1.0e-9_wp*x
      end do
   end do
   call output_vtk(grid, step=7)
   call output_npy(grid, step=7)
   call output_gnuplot(grid, step=7)
   call dealloc_grid(grid)
   print *, "outputs written"
end program
