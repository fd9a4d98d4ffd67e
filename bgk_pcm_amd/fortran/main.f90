! main.f90 -- bgk_thermal_gpu: the reference program's driver over the GPU
! step.
!
! Same input (input.dat in the working directory), same control flow and
! the same output files as the reference driver STEP_7/main.f90:
! initialization, VTK output at it=0 and every iter/noutput steps and at the
! end (Output_%08d.vtk), diagnostics every 1000 steps (fort.64, fort.666,
! fort.667, fort.668, fort.69, non-dim_bulk_numbers.dat, log.dat), Nusselt.dat
! at output steps, the time per step every 500 steps and the final Mlups
! line.  The time step runs on the GPU: between two steps where the
! reference prints or writes something, the lattice advances with one
! bgk_step call (the fused kernel for the no-slip cavity).
!
! Options (after the program name):
!   --stages   call the eight subroutines per step instead (BC_f, BC_g,
!              streaming, momentsPre, phase_field, forces, collision, moments)
!   --fp32     FP32 population storage
!   --mom      FP32 moment storage (rho and T in FP64; clean rounding)
!   --clean    FP64 temporaries where STEP_7 has implicit REAL(4) ones
!   --binary-vtk  write the split binary VTK files of STEP_CUF (Outphi_,
!              Outrho_, OutTem_, Outvel_) instead of the ASCII Output_ file
!   --dump-out write dumpFile.dat at the end of the run (restart: dump = 1
!              in input.dat reads it instead of the initialization)
!   --ext FILE physics extensions from the namelist &bgk_ext in FILE
!              (hot-wall temperature, conductivity ratio, cold east wall,
!              initial noise, critical and buoyancy reference temperatures;
!              all but the conductivity ratio need --mom)
program bgk_thermal_gpu
  use iso_c_binding
  use bgk_gpu
  implicit none

  type(bgk_params_t) :: p
  integer :: it0, iter, deltaOut, itOut, isignal, ev, s, k, nargs
  integer :: icountT0, icountT1, icount_rate, icount_max, icountL0, icountL1
  real(8) :: time_inn_loop, mem_mb, dev_mb
  logical :: per_stage, bin_vtk, dump_at_end
  character(len=256) :: arg, ext_file

  write(6,*) "===================================="
  write(6,*) " MI355X GPU version (bgk_pcm_amd)  "
  write(6,*) "===================================="

  call read_input('input.dat', p)
  per_stage = .false.
  bin_vtk = .false.
  dump_at_end = .false.
  ext_file = ''
  nargs = command_argument_count()
  k = 0
  do while (k < nargs)
     k = k + 1
     call get_command_argument(k, arg)
     select case (trim(arg))
     case ('--ext')
        if (k == nargs) then
           write(0,*) 'bgk_thermal_gpu: --ext needs a file name'
           stop 2
        end if
        k = k + 1
        call get_command_argument(k, ext_file)
     case ('--stages')
        per_stage = .true.
     case ('--fp32')
        p%precision = 1
     case ('--mom')
        p%precision = 2
        p%rounding = 1
     case ('--clean')
        p%rounding = 1
     case ('--binary-vtk')
        bin_vtk = .true.
     case ('--dump-out')
        dump_at_end = .true.
     case default
        write(0,*) 'bgk_thermal_gpu: unknown option ', trim(arg)
        stop 2
     end select
  end do
  call gpu_setup(p)
  if (len_trim(ext_file) > 0) call set_extensions(trim(ext_file))
  iter = p%iter
  isignal = 500

  if (p%dump == 0) then
     it = 0
     it0 = 1
     call initialization
     open(unit=150, file='log.dat')
     write(150,*) 'iteration', '    ', 'mass'
     write(150,*) '===================='
     close(150)
  else
     call dump_in
     it0 = it + 1
  end if

  deltaOut = int(iter / p%noutput)
  if (it == 0) then
     itOut = deltaOut
     if (bin_vtk) then
        call out2d_bin
     else
        call out2d
     end if
  else
     itOut = it + deltaOut
  end if

  open(unit=164, file='non-dim_bulk_numbers.dat', status='unknown')
  open(unit=165, file='Temp_trend.dat', status='unknown')

  call memory_report(mem_mb, dev_mb)
  write(6,1007) mem_mb
  write(6,1008) dev_mb
  call SYSTEM_CLOCK(icountL0, icount_rate, icount_max)
  icountT0 = icountL0
  if (it0 <= iter) write(*,*) "starting at ", it0

  it = it0
  do while (it <= iter)
     ! next step whose end the reference observes
     ev = it
     do while (ev < iter .and. mod(ev, 1000) /= 0 .and. ev /= itOut .and. mod(ev, isignal) /= 0)
        ev = ev + 1
     end do
     if (per_stage) then
        do s = it, ev
           call BC_f
           call BC_g
           call streaming
           call momentsPre
           call phase_field
           call forces
           call collision
           call moments
        end do
     else
        call steps(ev - it + 1)
     end if
     it = ev

     if (mod(it, 1000) == 0) then
        call diag
        call probe
        call prof_x
        call prof_y
        call field_check
        open(unit=150, file='log.dat', position='append')
        write(150, "(I8,A,F10.4)") it, ' ', mass
        close(150)
     end if

     if (it == itOut .or. it == iter) then
        if (bin_vtk) then
           call out2d_bin
        else
           call out2d
        end if
        open(unit=160, file='Nusselt.dat')
        do k = 1, 40
           write(160, "(I8,A,F10.4,A,F10.4,A,F10.4)") k, ' ', 0.0d0, ' ', 0.0d0, ' ', 0.0d0
        end do
        close(160)
        itOut = itOut + deltaOut
     end if

     if (mod(it, isignal) == 0) then
        call synchronize()
        call SYSTEM_CLOCK(icountL1, icount_rate, icount_max)
        time_inn_loop = real(icountL1 - icountL0) / (icount_rate)
        write(6,1003) (time_inn_loop) / isignal, it, iter
        call SYSTEM_CLOCK(icountL0, icount_rate, icount_max)
     end if
     it = it + 1
  end do

  call synchronize()
  call SYSTEM_CLOCK(icountT1, icount_rate, icount_max)
  time_inn_loop = real(icountT1 - icountT0) / (icount_rate)
  it = iter
  if (dump_at_end) call dump_out
  call memory_report(mem_mb, dev_mb)
  write(6,*) "=========================================="
  write(6,*) "INFO: Run completed"
  write(6,1004) (time_inn_loop)
  write(6,1005) (time_inn_loop) / iter
  write(6,1006) (float(p%nx) / 1000.0 * float(p%ny) / 1000.0) * (float(iter) / time_inn_loop)
  write(6,1007) mem_mb
  write(6,*) "=========================================="
  call gpu_finish()

1003 format("INFO: Mean   time           ",1(e14.6,1x),i8,"/",i8)
1004 format("INFO: Main loop time        ",1(e14.6,1x))
1005 format("INFO: Mean time for timestep",1(e14.6,1x))
1006 format("INFO: Mlups                 ",1(f14.6,1x))
1007 format("INFO: Memory used (MB)      ",1(f14.6,1x))
1008 format("INFO: Device memory (MB)    ",1(f14.6,1x))
end program bgk_thermal_gpu
