! Executable of the Fortran host (the reference's program sfincs, sfincs.f90:1-39):
! initialize, run to the stop time, finalize; exit status is the finalize status.
program sfincs_mi355x
   use sfincs_lib
   implicit none
   integer :: ierr, ierr2
   double precision :: deltat
   deltat = -999.9d0
   ierr = sfincs_initialize()
   if (ierr == 0) ierr = sfincs_update(deltat)
   ierr2 = sfincs_finalize()
   if (ierr /= 0) ierr2 = ierr
   call exit(ierr2)
end program
