! ref_bench.f90 -- times the reference STEP_OPENMP time step.
!
! TEST/BENCH INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) against the
! reference's own STEP_OPENMP sources, compiled where they lie with OpenMP,
! so the CPU baseline is the reference's own code.  It runs the eight calls
! of the reference time loop (STEP_OPENMP/main.f90: BC_f, BC_g, streaming,
! momentsPre, phase_field, forces, collision, moments) NSTEPS times without
! the reference's output and diagnostic calls, and prints one line:
!   nx=.. ny=.. steps=.. threads=.. seconds=.. mlups=.. rho=.. T=.. phi=..
! (the last three are interior means, a check that the run did the work).
! Usage (cwd must hold input.dat): ref_bench NSTEPS
program ref_bench
  use shared
  use omp_lib
  implicit none
  integer :: nsteps, s, nthreads
  real(8) :: tw0, tw1, mean_rho, mean_t, mean_phi
  character(len=64) :: arg

  call get_command_argument(1, arg)
  read(arg, *) nsteps
  call setup
  it = 0
  call initialization
  nthreads = omp_get_max_threads()
  tw0 = omp_get_wtime()
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
  tw1 = omp_get_wtime()
  mean_rho = sum(rho(1:Nx,1:Ny)) / (dble(Nx) * dble(Ny))
  mean_t = sum(T(1:Nx,1:Ny)) / (dble(Nx) * dble(Ny))
  mean_phi = sum(phi(1:Nx,1:Ny)) / (dble(Nx) * dble(Ny))
  write(6, '(A,I0,A,I0,A,I0,A,I0,A,ES16.8,A,F14.6,A,ES22.14,A,ES22.14,A,ES22.14)') &
       'nx=', Nx, ' ny=', Ny, ' steps=', nsteps, ' threads=', nthreads, &
       ' seconds=', tw1 - tw0, ' mlups=', dble(Nx) * dble(Ny) * dble(nsteps) / (tw1 - tw0) / 1.0d6, &
       ' rho=', mean_rho, ' T=', mean_t, ' phi=', mean_phi
end program ref_bench
