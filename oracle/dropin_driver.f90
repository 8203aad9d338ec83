! dropin_driver.f90 -- runs the product's drop-in newsolv_mg
! (hydro3d_amd/fortran/newsolv_mg.f90 -> newsolv_mg_h3d.f90, the GPU path)
! on module state set up exactly as ref_multiblock sets it up for the
! reference's loop, so that a test can compare the two bitwise.
!
! TEST INFRASTRUCTURE: linked with the reference objects that oracle/Makefile
! builds from src_v70 (modules, halo exchanges, press.for, alloc_dom.for),
! with the drop-in's objects and libh3dmg.so, and with dropin_stubs.f90
! (boundu/v/w as no-ops: the velocity boundary conditions of the host LES,
! which the reference's loop replay in ref_multiblock does not call either).
!
! Job (mb_setup's namelist, mode 'newsolv'): per block u*, v*, w* and p0 in;
! out: iter, rmax, then per block p (one ghost layer) and su, then per block
! the whole u, v, w arrays.
program dropin_driver
  use mb_setup
  implicit none

  call start()
  if (trim(mode) /= 'newsolv') then
     print *, 'dropin_driver: mode must be newsolv'
     call MPI_ABORT(MPI_COMM_WORLD, 2, ierr)
  end if
  numfile2 = 1002
  open(unit=numfile2, file='worktime.dat')
  call read_velocity_state()
  call newsolv_mg
  write(uout) dble(iter), rmax
  call write_pressure_state()
  call write_velocity_arrays()
  close(numfile2)
  call finish()
end program dropin_driver
