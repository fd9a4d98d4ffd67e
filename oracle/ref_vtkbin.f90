! ref_vtkbin.f90 -- driver for the reference's output routines.
!
! TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) against the
! reference's STEP_7 sources and the split binary VTK writers of STEP_CUF
! (out2d.phi.f90, out2d.rho.f90, out2d.temp.f90, compiled where they lie
! against STEP_7's module shared), to write the files that pin this
! library's binary VTK and restart writers (tests/golden/).
!
! Usage (cwd must hold input.dat, read by the reference's own setup):
!   ref_vtkbin NSTEPS
! runs NSTEPS time steps (STEP_7/main.f90:60-69), then writes
! Outphi_/Outrho_/OutTem_<it>.vtk (STEP_CUF/out2d.*.f90) and dumpFile.dat
! (STEP_7/dump_out.f90).
program ref_vtkbin
  use shared
  implicit none
  integer :: nsteps, s
  character(len=64) :: arg

  call get_command_argument(1, arg)
  read(arg, *) nsteps
  call setup
  it = 0
  call initialization
  do s = 1, nsteps
     it = s
     call BC_f
     call BC_g
     call streaming
     call momentsPre
     call phase_field
     call forces
     call collision
     call moments
  end do
  call out2d_phi
  call out2d_rho
  call out2d_temp
  call dump_out
end program ref_vtkbin
