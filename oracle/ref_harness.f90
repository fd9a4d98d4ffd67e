! ref_harness.f90 -- driver for the reference STEP_7 subroutines.
!
! TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) against the
! reference's own STEP_7 sources, compiled where they lie, to write binary
! snapshots that pin the C restatement (oracle/bgk_oracle.c) and the golden
! fixtures under tests/golden/.  Nothing here ships with the library.
!
! Usage (cwd must hold input.dat, read by the reference's own setup):
!   ref_harness NSTEPS STAGE_FIRST STAGE_LAST OUTDIR [DIAG_EVERY]
! writes OUTDIR/init.bin after initialization, OUTDIR/sNNNN_K.bin before
! (K=0) and after stage K (1..8) of every step NNNN in
! STAGE_FIRST..STAGE_LAST, and OUTDIR/final.bin after NSTEPS steps.  Each .bin holds
! the 46 fields in the order of include/bgk_pcm.h (f0..f8, fp0..fp8,
! g0..g8, gp0..gp8, rho, u, v, T, rho2, u2, v2, T2, phi, phi_prec), each a
! (0:Nx+1,0:Ny+1) real(mykind) array, raw little-endian stream.  Every
! DIAG_EVERY steps (default 1000) it calls diag, probe, prof_x, prof_y and
! field_check as the reference time loop does (STEP_7/main.f90:72-82).
program ref_harness
  use shared
  implicit none
  integer :: nsteps, sfirst, slast, diag_every, s, nargs
  character(len=512) :: arg, outdir

  nargs = command_argument_count()
  if (nargs < 4) then
     write(0,*) 'usage: ref_harness NSTEPS STAGE_FIRST STAGE_LAST OUTDIR [DIAG_EVERY]'
     stop 2
  end if
  call get_command_argument(1, arg); read(arg,*) nsteps
  call get_command_argument(2, arg); read(arg,*) sfirst
  call get_command_argument(3, arg); read(arg,*) slast
  call get_command_argument(4, outdir)
  diag_every = 1000
  if (nargs >= 5) then
     call get_command_argument(5, arg); read(arg,*) diag_every
  end if

  call setup
  it = 0
  call initialization
  open(unit=164, file='non-dim_bulk_numbers.dat', status='unknown')
  call snapshot(trim(outdir)//'/init.bin')

  do s = 1, nsteps
     it = s
     if (win(s)) call snapshot(stage_name(s, 0))
     call BC_f
     if (win(s)) call snapshot(stage_name(s, 1))
     call BC_g
     if (win(s)) call snapshot(stage_name(s, 2))
     call streaming
     if (win(s)) call snapshot(stage_name(s, 3))
     call momentsPre
     if (win(s)) call snapshot(stage_name(s, 4))
     call phase_field
     if (win(s)) call snapshot(stage_name(s, 5))
     call forces
     if (win(s)) call snapshot(stage_name(s, 6))
     call collision
     if (win(s)) call snapshot(stage_name(s, 7))
     call moments
     if (win(s)) call snapshot(stage_name(s, 8))
     if (diag_every > 0) then
        if (mod(it, diag_every) == 0) then
           call diag
           call probe
           call prof_x
           call prof_y
           call field_check
        end if
     end if
  end do
  call snapshot(trim(outdir)//'/final.bin')
  close(164)

contains

  logical function win(step)
    integer, intent(in) :: step
    win = (step >= sfirst .and. step <= slast)
  end function win

  function stage_name(step, k) result(name)
    integer, intent(in) :: step, k
    character(len=600) :: name
    write(name, '(A,A,I4.4,A,I1,A)') trim(outdir), '/s', step, '_', k, '.bin'
  end function stage_name

  subroutine snapshot(path)
    character(len=*), intent(in) :: path
    integer :: un
    open(newunit=un, file=trim(path), access='stream', form='unformatted', status='replace')
    write(un) f0, f1, f2, f3, f4, f5, f6, f7, f8
    write(un) fp0, fp1, fp2, fp3, fp4, fp5, fp6, fp7, fp8
    write(un) g0, g1, g2, g3, g4, g5, g6, g7, g8
    write(un) gp0, gp1, gp2, gp3, gp4, gp5, gp6, gp7, gp8
    write(un) rho, u, v, T, rho2, u2, v2, T2, phi, phi_prec
    close(un)
  end subroutine snapshot

end program ref_harness
