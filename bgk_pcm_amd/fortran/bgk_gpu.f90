! bgk_gpu.f90 -- Fortran side of the drop-in boundary.
!
! ISO_C_BINDING interfaces to libbgk_pcm (include/bgk_pcm.h) and the
! subroutines a Fortran time loop calls, with the reference's names and
! meaning (STEP_7/main.f90:60-105): initialization, BC_f, BC_g, streaming,
! momentsPre, phase_field, forces, collision, moments, diag, probe, prof_x,
! prof_y, field_check, mass_check, out2d.  The lattice lives on the GPU; the
! diagnostics fetch what they print (reductions or gathers) and write the
! reference's files with the reference's formats.  Every failing call stops
! the program with the library's message.
module bgk_gpu
  use iso_c_binding
  implicit none

  type, bind(C) :: bgk_params_t
     integer(c_int32_t) :: iter, noutput, dump, nx, ny, reserved0
     real(c_double) :: re, rho0, rho1, tau_f, tau_g, t0, t_in_sol, grav, eps, &
                       stefan_number, t_width, length
     integer(c_int32_t) :: west_bc, east_bc, north_bc, south_bc
     integer(c_int32_t) :: precision, rounding, device, fp32_collision
  end type bgk_params_t

  ! physics extensions (bgk_ext; zeros and k_ratio = 1 keep STEP_7)
  type, bind(C) :: bgk_ext_t
     real(c_double) :: t_hot, k_ratio
     integer(c_int32_t) :: east_cold, reserved
     real(c_double) :: t_cold, noise_amp
     integer(c_int64_t) :: noise_seed
     real(c_double) :: t_crit, t_ref
  end type bgk_ext_t

  type, bind(C) :: bgk_info_t
     integer(c_int64_t) :: device_bytes, device_free, device_total
     real(c_double) :: host_maxrss_mb
     integer(c_int32_t) :: pitch, fused_path
  end type bgk_info_t

  ! field ids of include/bgk_pcm.h
  integer(c_int32_t), parameter :: F_F0 = 0, F_G0 = 18, F_RHO = 36, F_U = 37, F_V = 38, &
                                   F_T = 39, F_PHI = 44

  interface
     integer(c_int) function c_bgk_create(p, h) bind(C, name='bgk_create')
       import :: bgk_params_t, c_ptr, c_int
       type(bgk_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function
     integer(c_int) function c_bgk_destroy(h) bind(C, name='bgk_destroy')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     type(c_ptr) function c_bgk_last_error() bind(C, name='bgk_last_error')
       import :: c_ptr
     end function
     integer(c_int) function c_bgk_initialization(h) bind(C, name='bgk_initialization')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_bc_f(h) bind(C, name='bgk_bc_f')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_bc_g(h) bind(C, name='bgk_bc_g')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_streaming(h) bind(C, name='bgk_streaming')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_moments_pre(h) bind(C, name='bgk_moments_pre')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_phase_field(h) bind(C, name='bgk_phase_field')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_forces(h) bind(C, name='bgk_forces')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_collision(h) bind(C, name='bgk_collision')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_moments(h) bind(C, name='bgk_moments')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_step(h, n) bind(C, name='bgk_step')
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: n
     end function
     integer(c_int) function c_bgk_synchronize(h) bind(C, name='bgk_synchronize')
       import :: c_ptr, c_int
       type(c_ptr), value :: h
     end function
     integer(c_int) function c_bgk_download(h, field, dst, ld) bind(C, name='bgk_download')
       import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: field
       real(c_double), intent(out) :: dst(*)
       integer(c_int64_t), value :: ld
     end function
     integer(c_int) function c_bgk_upload(h, field, src, ld) bind(C, name='bgk_upload')
       import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: field
       real(c_double), intent(in) :: src(*)
       integer(c_int64_t), value :: ld
     end function
     integer(c_int) function c_bgk_field_check(h, out) bind(C, name='bgk_field_check')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(3)
     end function
     integer(c_int) function c_bgk_mass_check(h, out) bind(C, name='bgk_mass_check')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(3)
     end function
     integer(c_int) function c_bgk_diag_sums(h, out) bind(C, name='bgk_diag_sums')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(4)
     end function
     integer(c_int) function c_bgk_gather(h, field, n, i, j, out) bind(C, name='bgk_gather')
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: field, n
       integer(c_int32_t), intent(in) :: i(*), j(*)
       real(c_double), intent(out) :: out(*)
     end function
     integer(c_int) function c_bgk_set_extensions(h, e) bind(C, name='bgk_set_extensions')
       import :: c_ptr, c_int, bgk_ext_t
       type(c_ptr), value :: h
       type(bgk_ext_t), intent(in) :: e
     end function
     integer(c_int) function c_bgk_get_info(h, info) bind(C, name='bgk_get_info')
       import :: c_ptr, c_int, bgk_info_t
       type(c_ptr), value :: h
       type(bgk_info_t), intent(out) :: info
     end function
  end interface

  type(c_ptr), save :: handle = c_null_ptr
  type(bgk_params_t), save :: prm
  real(c_double), save :: kvisc, alpha, tin
  real(c_double), save :: mass = 0.0d0
  ! the time step counter of the driver loop (module shared's `it`,
  ! STEP_7/shared.f90), read by the argument-less diagnostics below as the
  ! reference's subroutines read it (STEP_7/main.f90:72-82, 90)
  integer, save :: it = 0

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc == 0) return
    call c_f_pointer(c_bgk_last_error(), msg, [1024])
    write(0, '(A,A,I0,A)', advance='no') 'bgk_gpu: ', what//' failed (', rc, '): '
    do k = 1, 1024
       if (msg(k) == c_null_char) exit
       write(0, '(A)', advance='no') msg(k)
    end do
    write(0, *)
    stop 1
  end subroutine check

  ! input.dat: 21 list-directed records "label value" (STEP_7/setup.f90:7-31)
  subroutine read_input(path, p)
    character(len=*), intent(in) :: path
    type(bgk_params_t), intent(out) :: p
    character :: lbl
    integer :: u, iv(10)
    real(8) :: rv(12)
    integer :: k
    open(newunit=u, file=path, status='old', action='read')
    read(u, *) lbl, iv(1)
    read(u, *) lbl, iv(2)
    read(u, *) lbl, iv(3)
    read(u, *) lbl, iv(4)
    read(u, *) lbl, iv(5)
    do k = 1, 12
       read(u, *) lbl, rv(k)
    end do
    do k = 6, 9
       read(u, *) lbl, iv(k)
    end do
    close(u)
    p%iter = iv(1); p%noutput = iv(2); p%dump = iv(3); p%nx = iv(4); p%ny = iv(5)
    p%reserved0 = 0
    p%re = rv(1); p%rho0 = rv(2); p%rho1 = rv(3); p%tau_f = rv(4); p%tau_g = rv(5)
    p%t0 = rv(6); p%t_in_sol = rv(7); p%grav = rv(8); p%eps = rv(9)
    p%stefan_number = rv(10); p%t_width = rv(11); p%length = rv(12)
    p%west_bc = iv(6); p%east_bc = iv(7); p%north_bc = iv(8); p%south_bc = iv(9)
    p%precision = 0; p%rounding = 0; p%device = 0; p%fp32_collision = 0
  end subroutine read_input

  ! allocation and constants on the device (STEP_7/setup.f90:34-115)
  subroutine gpu_setup(p)
    type(bgk_params_t), intent(in) :: p
    prm = p
    call check(c_bgk_create(prm, handle), 'bgk_create')
    kvisc = (1.d0/3.d0) * (prm%tau_f - 0.5d0)
    alpha = (1.d0/3.d0) * (prm%tau_g - 0.5d0)
    tin = prm%t0
  end subroutine gpu_setup

  ! Physics extensions from a namelist file, before the initialization:
  !   &bgk_ext t_hot = 283.5, k_ratio = 2.0, t_cold = 273.0,
  !            noise_amp = 1.0d-3, noise_seed = 7, t_crit = 273.0, t_ref = 0 /
  ! (any subset; t_cold > 0 puts the east wall at t_cold).  t_hot also
  ! becomes the west-wall temperature of the diagnostics.
  subroutine set_extensions(path)
    character(len=*), intent(in) :: path
    real(8) :: t_hot, k_ratio, t_cold, noise_amp, t_crit, t_ref
    integer(8) :: noise_seed
    type(bgk_ext_t) :: e
    integer :: u
    namelist /bgk_ext/ t_hot, k_ratio, t_cold, noise_amp, noise_seed, t_crit, t_ref
    t_hot = 0.d0; k_ratio = 1.d0; t_cold = 0.d0; noise_amp = 0.d0; noise_seed = 0
    t_crit = 0.d0; t_ref = 0.d0
    open(newunit=u, file=path, status='old', action='read')
    read(u, nml=bgk_ext)
    close(u)
    e%t_hot = t_hot; e%k_ratio = k_ratio
    e%east_cold = merge(1, 0, t_cold > 0.d0); e%reserved = 0; e%t_cold = t_cold
    e%noise_amp = noise_amp; e%noise_seed = noise_seed
    e%t_crit = t_crit; e%t_ref = t_ref
    call check(c_bgk_set_extensions(handle, e), 'bgk_set_extensions')
    if (t_hot > 0.d0) tin = t_hot
  end subroutine set_extensions

  subroutine gpu_finish()
    if (c_associated(handle)) call check(c_bgk_destroy(handle), 'bgk_destroy')
    handle = c_null_ptr
  end subroutine gpu_finish

  subroutine initialization()
    real(c_double) :: u0
    call check(c_bgk_initialization(handle), 'initialization')
    u0 = prm%re * kvisc / prm%length
    write(*,*) '****************************************************'
    if (prm%precision == 0) then
       write(*,*) 'Double precision used (GPU)                         '
    else if (prm%precision == 1) then
       write(*,*) 'FP32 population storage used (GPU)                  '
    else
       write(*,*) 'FP32 moment storage used (GPU)                      '
    end if
    write(*,*) '****************************************************'
    write(*,*) 'Nx   = ', prm%nx
    write(*,*) 'Ny   = ', prm%ny
    write(*,*) 'alpha= ', alpha
    write(*,*) 'kvisc= ', kvisc
    write(*,*) 'u0   = ', u0
    write(*,*) 'T0   = ', prm%t0
    write(*,*) '****************************************************'
    if (u0 >= sqrt(1.d0/3.d0)/3.d0) then
       write(*,*) '=============================='
       write(*,*) 'WARNING! Mach number too high!'
       write(*,*) '=============================='
    end if
  end subroutine initialization

  subroutine BC_f()
    call check(c_bgk_bc_f(handle), 'BC_f')
  end subroutine BC_f
  subroutine BC_g()
    call check(c_bgk_bc_g(handle), 'BC_g')
  end subroutine BC_g
  subroutine streaming()
    call check(c_bgk_streaming(handle), 'streaming')
  end subroutine streaming
  subroutine momentsPre()
    call check(c_bgk_moments_pre(handle), 'momentsPre')
  end subroutine momentsPre
  subroutine phase_field()
    call check(c_bgk_phase_field(handle), 'phase_field')
  end subroutine phase_field
  subroutine forces()
    call check(c_bgk_forces(handle), 'forces')
  end subroutine forces
  subroutine collision()
    call check(c_bgk_collision(handle), 'collision')
  end subroutine collision
  subroutine moments()
    call check(c_bgk_moments(handle), 'moments')
  end subroutine moments

  ! n consecutive time steps (the fused kernel for the no-slip cavity)
  subroutine steps(n)
    integer, intent(in) :: n
    call check(c_bgk_step(handle, int(n, c_int32_t)), 'bgk_step')
  end subroutine steps

  ! host memory high-water mark in MB (STEP_7/get_mem.c:23-38) and the
  ! device bytes this lattice holds
  subroutine memory_report(host_mb, device_mb)
    real(8), intent(out) :: host_mb, device_mb
    type(bgk_info_t) :: info
    call check(c_bgk_get_info(handle, info), 'bgk_get_info')
    host_mb = info%host_maxrss_mb
    device_mb = dble(info%device_bytes) / 1048576.d0
  end subroutine memory_report

  subroutine synchronize()
    call check(c_bgk_synchronize(handle), 'bgk_synchronize')
  end subroutine synchronize

  subroutine download(field, a)
    integer(c_int32_t), intent(in) :: field
    real(c_double), intent(out) :: a(0:prm%nx+1, 0:prm%ny+1)
    call check(c_bgk_download(handle, field, a, int(prm%nx + 2, c_int64_t)), 'bgk_download')
  end subroutine download

  subroutine upload(field, a)
    integer(c_int32_t), intent(in) :: field
    real(c_double), intent(in) :: a(0:prm%nx+1, 0:prm%ny+1)
    call check(c_bgk_upload(handle, field, a, int(prm%nx + 2, c_int64_t)), 'bgk_upload')
  end subroutine upload

  ! values of u, v, rho, T, phi at n nodes
  subroutine gather5(n, ii, jj, vals)
    integer, intent(in) :: n
    integer(c_int32_t), intent(in) :: ii(n), jj(n)
    real(c_double), intent(out) :: vals(n, 5)
    integer(c_int32_t), parameter :: ids(5) = [F_U, F_V, F_RHO, F_T, F_PHI]
    integer :: k
    do k = 1, 5
       call check(c_bgk_gather(handle, ids(k), int(n, c_int32_t), ii, jj, vals(:, k)), 'bgk_gather')
    end do
  end subroutine gather5

  ! STEP_7/field_check.f90:11-24 -> fort.64
  subroutine field_check()
    real(c_double) :: m(3)
    call check(c_bgk_field_check(handle, m), 'bgk_field_check')
    write(64, 1002) it, m(1), m(2), m(3)
    flush(64)
1002 format(i8,1x,3(e16.8,1x))
  end subroutine field_check

  ! STEP_7/mass_check.f90:8-32 -> fort.64 and stdout; sets mass
  subroutine mass_check()
    real(c_double) :: m(3)
    call check(c_bgk_mass_check(handle, m), 'bgk_mass_check')
    mass = m(1)
    write(64, *) it, m(2), m(3)
    write(6, 1001) m(2), m(3)
1001 format("CHECK: Mass_aver, Temp_aver",2(e14.6,1x))
  end subroutine mass_check

  ! STEP_7/probe.f90:7-15 -> fort.666
  subroutine probe()
    integer(c_int32_t) :: ii(1), jj(1)
    real(c_double) :: vals(1, 5)
    ii(1) = prm%nx / 2
    jj(1) = prm%ny / 2
    call gather5(1, ii, jj, vals)
    write(666, 1001) it, vals(1, 1), vals(1, 2), vals(1, 3), vals(1, 4), vals(1, 5)
1001 format(i8,5(e14.6,1x))
  end subroutine probe

  ! STEP_7/prof_x.f90:7-17 -> fort.667
  subroutine prof_x()
    integer(c_int32_t), allocatable :: ii(:), jj(:)
    real(c_double), allocatable :: vals(:, :)
    integer :: i
    allocate(ii(prm%nx), jj(prm%nx), vals(prm%nx, 5))
    do i = 1, prm%nx
       ii(i) = i
       jj(i) = prm%ny / 2
    end do
    call gather5(prm%nx, ii, jj, vals)
    write(667, 1005) it
    do i = 1, prm%nx
       write(667, 1002) i, vals(i, 1), vals(i, 2), vals(i, 3), vals(i, 4), vals(i, 5)
    end do
    write(667, '(a1)')
    write(667, '(a1)')
1002 format(i8,5(e14.6,1x))
1005 format("# t=",i7)
  end subroutine prof_x

  ! STEP_7/prof_y.f90:7-17 -> fort.668
  subroutine prof_y()
    integer(c_int32_t), allocatable :: ii(:), jj(:)
    real(c_double), allocatable :: vals(:, :)
    integer :: j
    allocate(ii(prm%ny), jj(prm%ny), vals(prm%ny, 5))
    do j = 1, prm%ny
       ii(j) = prm%nx / 2
       jj(j) = j
    end do
    call gather5(prm%ny, ii, jj, vals)
    write(668, 1005) it
    do j = 1, prm%ny
       write(668, 1002) j, vals(j, 1), vals(j, 2), vals(j, 3), vals(j, 4), vals(j, 5)
    end do
    write(668, '(a1)')
    write(668, '(a1)')
1002 format(i8,5(e14.6,1x))
1005 format("# t=",i7)
  end subroutine prof_y

  ! STEP_7/diag.f90:15-131: non-dimensional numbers from the device sums
  ! -> fort.69, non-dim_bulk_numbers.dat (unit 164, opened by the driver)
  subroutine diag()
    real(c_double) :: s(4)
    real(8) :: beta_gras, dT, gras, pr, ray, fou, theta, theta1, theta2, nus, &
               nus_comp, nus_comp_2, melt, front, front_exp
    integer :: nx, ny
    nx = prm%nx
    ny = prm%ny
    call check(c_bgk_diag_sums(handle, s), 'bgk_diag_sums')
    dT = prm%t0 - prm%t_in_sol
    beta_gras = (prm%rho0 / prm%rho1 - 1) / dT
    gras = prm%grav * beta_gras * dT * ny**3 / kvisc**2
    pr = (2. * prm%tau_f - 1.) / (2. * prm%tau_g - 1.)
    ray = pr * gras
    fou = alpha * it / (ny**2)
    theta = prm%stefan_number * fou
    theta1 = ray**(-0.5d0)
    theta2 = (nx / ny) * ray**(-1./4.)
    write(69, *) 'beta_gras, grav, Gras, Delta_Temp, Pr, Ray, Fou, theta'
    write(69, *) beta_gras, prm%grav, gras, dT, pr, ray, fou, theta
    write(69, *) '---------------------------------------------'
    flush(69)
    nus = ((2*theta)**(-0.5d0)) + (0.35d0 * (ray**(0.25d0)) - ((2*theta)**(-0.5d0))) * &
          (1. + (0.0175d0 * (ray**(3./4.)) * (theta**(3./2.)))**(-2.d0))**(1./(-2.))
    nus_comp = s(1) / ny
    nus_comp_2 = s(2) / ny
    front = 0.0d0
    front_exp = 0.0d0
    if (mod(it, 100) == 0) then
       melt = s(3) / (nx * ny)
       if (theta <= theta1) then
          front = ny / ((2*theta)**(-0.5d0))
          front_exp = ny / nus
       else if (theta > theta1 .and. theta < theta2) then
          front = s(4) / ny
          front_exp = ny * theta * ray**(1./4.)
       end if
       write(6, *) 'CHECK: theta =', theta
       write(164, 640) it, theta, nus, nus_comp, nus_comp_2, melt, front, front_exp, &
                       gras, pr, ray, fou
    end if
640 format(1I8,11(1x,e13.6))
  end subroutine diag

  ! STEP_7/out2d.f90:8-65: legacy ASCII VTK of rho, T, phi, (u,v,0)
  subroutine out2d()
    character(len=25) :: fname
    real(c_double), allocatable :: a(:, :), b(:, :)
    integer :: un, i, l
    allocate(a(0:prm%nx+1, 0:prm%ny+1), b(0:prm%nx+1, 0:prm%ny+1))
    fname = 'Output_xxxxxxxx.vtk'
    write(fname(8:15), '(i8.8)') it
    open(newunit=un, file=fname, status='unknown')
    write(un, '(A)') '# vtk DataFile Version 2.0'
    write(un, '(A)') 'Griglia strutturata'
    write(un, '(A)') 'ASCII'
    write(un, '(A)') 'DATASET STRUCTURED_GRID'
    write(un, '(A,3I8)') 'DIMENSIONS', prm%nx + 2, prm%ny + 2, 1
    write(un, '(A,I15,A)') 'POINTS', (prm%nx + 2) * (prm%ny + 2) * 1, ' float'
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) i, l, 1
       end do
    end do
    write(un, '(A,I15)') 'POINT_DATA', (prm%nx + 2) * (prm%ny + 2) * 1
    write(un, '(A)') 'SCALARS DENSITY float 1'
    write(un, '(A)') 'LOOKUP_TABLE default'
    call download(F_RHO, a)
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l)
       end do
    end do
    write(un, '(A)') 'SCALARS Temperature float 1'
    write(un, '(A)') 'LOOKUP_TABLE default'
    call download(F_T, a)
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l)
       end do
    end do
    write(un, '(A)') 'SCALARS Phase_Field float 1'
    write(un, '(A)') 'LOOKUP_TABLE default'
    call download(F_PHI, a)
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l)
       end do
    end do
    write(un, '(A)') 'VECTORS VELOCITY float'
    call download(F_U, a)
    call download(F_V, b)
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l), b(i, l), 0.0d0
       end do
    end do
    close(un)
  end subroutine out2d

  ! Split binary VTK files of STEP_CUF (out2d.phi.f90, out2d.rho.f90,
  ! out2d.temp.f90, out2d.vel.f90): ASCII header, then the interior nodes
  ! (j outer, i inner) as big-endian real(8) stream records.
  subroutine out2d_bin()
    real(c_double), allocatable :: a(:, :), b(:, :)
    allocate(a(0:prm%nx+1, 0:prm%ny+1), b(0:prm%nx+1, 0:prm%ny+1))
    call download(F_PHI, a)
    call bin_scalar('Outphi_', it, 'SCALARS phi double', a)
    call download(F_RHO, a)
    call bin_scalar('Outrho_', it, 'SCALARS rho double', a)
    call download(F_T, a)
    call bin_scalar('OutTem_', it, 'SCALARS Temperature double', a)
    call download(F_U, a)
    call download(F_V, b)
    call bin_vel(it, a, b)
    deallocate(a, b)
  end subroutine out2d_bin

  subroutine bin_header(un, fname, it, title)
    integer, intent(out) :: un
    character(len=25), intent(inout) :: fname
    integer, intent(in) :: it
    character(len=*), intent(in) :: title
    write(fname(8:15), '(i8.8)') it
    open(newunit=un, file=fname, status='unknown')
    write(un, '(A)') '# vtk DataFile Version 2.0'
    write(un, '(A)') 'Campo'
    write(un, '(A)') 'BINARY'
    write(un, '(A)') 'DATASET STRUCTURED_POINTS'
    write(un, '(A,3I8)') 'DIMENSIONS', prm%nx, prm%ny, 1
    write(un, '(A7,I10,A1,I10,A1,I10)') 'ORIGIN ', 1, ' ', 1, ' ', 1
    write(un, '(A8,I10,A1,I10,A1,I10)') 'SPACING ', 1, ' ', 1, ' ', 1
    write(un, '(A10,I10)') 'POINT_DATA ', prm%nx * prm%ny
    write(un, '(A)') title
  end subroutine bin_header

  subroutine bin_scalar(prefix, it, title, a)
    character(len=*), intent(in) :: prefix, title
    integer, intent(in) :: it
    real(c_double), intent(in) :: a(0:prm%nx+1, 0:prm%ny+1)
    character(len=25) :: fname
    integer :: un, i, l
    fname = prefix // 'xxxxxxxx.vtk'
    call bin_header(un, fname, it, title)
    write(un, '(A)') 'LOOKUP_TABLE default'
    close(un)
    open(newunit=un, file=fname, status='old', position='append', form='unformatted', &
         access='stream', convert='BIG_ENDIAN')
    do l = 1, prm%ny
       do i = 1, prm%nx
          write(un) a(i, l)
       end do
    end do
    close(un)
  end subroutine bin_scalar

  subroutine bin_vel(it, a, b)
    integer, intent(in) :: it
    real(c_double), intent(in) :: a(0:prm%nx+1, 0:prm%ny+1), b(0:prm%nx+1, 0:prm%ny+1)
    character(len=25) :: fname
    integer :: un, i, l
    fname = 'Outvel_xxxxxxxx.vtk'
    call bin_header(un, fname, it, 'VECTORS velocity double')
    close(un)
    open(newunit=un, file=fname, status='old', position='append', form='unformatted', &
         access='stream', convert='BIG_ENDIAN')
    do l = 1, prm%ny
       do i = 1, prm%nx
          write(un) a(i, l), b(i, l), 0.0d0
       end do
    end do
    close(un)
  end subroutine bin_vel

  ! Restart file: STEP_7/dump_out.f90's records (it; per node u, v, rho,
  ! f0..f8 and T, g0..g8, list-directed), followed by phi of every node,
  ! which the reference's restart does not keep (its dump_in restores no
  ! phi, so a run restarted from it melts from phi = 0).  A reader of the
  ! reference format reads the first part unchanged.
  subroutine dump_out()
    real(c_double), allocatable :: a(:, :, :)
    integer :: un, i, l, k
    allocate(a(0:prm%nx+1, 0:prm%ny+1, 23))
    call download(F_U, a(:, :, 1))
    call download(F_V, a(:, :, 2))
    call download(F_RHO, a(:, :, 3))
    do k = 0, 8
       call download(F_F0 + k, a(:, :, 4 + k))
       call download(F_G0 + k, a(:, :, 14 + k))
    end do
    call download(F_T, a(:, :, 13))
    call download(F_PHI, a(:, :, 23))
    open(newunit=un, file='dumpFile.dat')
    write(un, *) it
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l, 1), a(i, l, 2), a(i, l, 3), a(i, l, 4), a(i, l, 5), a(i, l, 6), &
                       a(i, l, 7), a(i, l, 8), a(i, l, 9), a(i, l, 10), a(i, l, 11), a(i, l, 12)
          write(un, *) a(i, l, 13), a(i, l, 14), a(i, l, 15), a(i, l, 16), a(i, l, 17), &
                       a(i, l, 18), a(i, l, 19), a(i, l, 20), a(i, l, 21), a(i, l, 22)
       end do
    end do
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          write(un, *) a(i, l, 23)
       end do
    end do
    close(un)
    deallocate(a)
  end subroutine dump_out

  ! Restart from dumpFile.dat (dump_out above; STEP_7/dump_in.f90 reads the
  ! same first part).  Stops when the phi block is missing.
  subroutine dump_in()
    real(c_double), allocatable :: a(:, :, :)
    integer :: un, i, l, k, ios
    allocate(a(0:prm%nx+1, 0:prm%ny+1, 23))
    open(newunit=un, file='dumpFile.dat', status='old')
    read(un, *) it
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          read(un, *) a(i, l, 1), a(i, l, 2), a(i, l, 3), a(i, l, 4), a(i, l, 5), a(i, l, 6), &
                      a(i, l, 7), a(i, l, 8), a(i, l, 9), a(i, l, 10), a(i, l, 11), a(i, l, 12)
          read(un, *) a(i, l, 13), a(i, l, 14), a(i, l, 15), a(i, l, 16), a(i, l, 17), &
                      a(i, l, 18), a(i, l, 19), a(i, l, 20), a(i, l, 21), a(i, l, 22)
       end do
    end do
    do l = 0, prm%ny + 1
       do i = 0, prm%nx + 1
          read(un, *, iostat=ios) a(i, l, 23)
          if (ios /= 0) then
             write(0, *) 'bgk_thermal_gpu: dumpFile.dat has no phi block (a restart without'
             write(0, *) 'the phase field would not continue the run)'
             stop 1
          end if
       end do
    end do
    close(un)
    call upload(F_U, a(:, :, 1))
    call upload(F_V, a(:, :, 2))
    call upload(F_RHO, a(:, :, 3))
    do k = 0, 8
       call upload(F_F0 + k, a(:, :, 4 + k))
       call upload(F_G0 + k, a(:, :, 14 + k))
    end do
    call upload(F_T, a(:, :, 13))
    call upload(F_PHI, a(:, :, 23))
    call upload(F_PHI + 1, a(:, :, 23))
    deallocate(a)
  end subroutine dump_in

end module bgk_gpu
