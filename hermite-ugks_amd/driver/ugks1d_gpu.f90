! ------------------------------------------------------------
! 1D Hermite-UGKS driver over libhugks.so.
!
! Usage: ugks1d_gpu [input.namelist]
!
! The input is the reference's &params namelist (1d/src/config.f90:47-49,
! same names and defaults) optionally followed by a &gpu group
! (precision = 'fp64' | 'mixed'; max_steps, 0 = no cap).  The driver keeps
! what the reference main program does around the step - case selection,
! velocity quadrature, initial condition, the maxtime and Fourier
! steady-state exits, macro<iter>.dat files (1d/src/main.f90:24-140) - and
! hands the step itself, including the steady-state test, to the library.
! ------------------------------------------------------------
module ugks1d_setup
  use, intrinsic :: iso_c_binding
  use hugks_capi
  use hugks_quadrature
  implicit none

  ! &params, defaults of 1d/src/config.f90:11-45
  integer :: nx = 600
  real(8) :: xlength = 1.0d0
  integer :: nv = 100
  real(8) :: vmax = 5.0d0, Tref = 1.0d0
  character(len=20) :: quad_type = 'newton_cotes'
  character(len=20) :: force_type = 'hermite'
  integer :: Nm = 6
  real(8) :: cfl = 0.5d0, maxtime = 0.2d0
  real(8) :: steady_tol = 1.d-9, mass_tol = 1.d-10
  integer :: steady_steps = 1000
  integer :: case_id = 3
  real(8) :: knudsen = 0.01d0, phi = -1.0d0, TwL = 1.0d0, TwR = 1.1d0
  character(len=40) :: refpfile = ''
  namelist /params/ nx, xlength, nv, vmax, Tref, quad_type, force_type, Nm, &
                    cfl, maxtime, steady_tol, mass_tol, steady_steps, case_id, &
                    knudsen, phi, TwL, TwR, refpfile

  character(len=16) :: precision = 'fp64'
  integer :: max_steps = 0
  character(len=16) :: rt_variant = 'reference'
  integer :: seed = 0
  namelist /gpu/ precision, max_steps, rt_variant, seed

  ! constants of the reference (1d/src/data1d.f90:9-10,63; Rc = 0.5)
  real(8), parameter :: PI = 3.1415926535d0
  real(8), parameter :: PRANDTL = 0.666666666667d0
  real(8), parameter :: RC = 0.5d0

  real(8), allocatable, target :: uspace(:), weight(:), prim0(:, :)
  real(8), allocatable :: xc(:)
  real(8) :: dx, gam

contains

  subroutine read_config(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios
    logical :: found
    inquire(file=fname, exist=found)
    if (.not. found) then
      write(*, *) "config file '", trim(fname), "' not found; using built-in defaults"
      return
    end if
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, *) "error opening '", trim(fname), "'; using defaults"
      return
    end if
    read(u, nml=params, iostat=ios)
    if (ios /= 0) write(*, *) "warning: namelist parse error in '", trim(fname), &
                              "'; using partial defaults"
    rewind(u)
    read(u, nml=gpu, iostat=ios)
    close(u)
  end subroutine read_config

  real(8) function get_mu(kn, alpha, omega)
    real(8), intent(in) :: kn, alpha, omega
    get_mu = 5.d0 * (alpha + 1.d0) * (alpha + 2.d0) * sqrt(PI) / &
             (4.d0 * alpha * (5.d0 - 2.d0 * omega) * (7.d0 - 2.d0 * omega)) * kn
  end function get_mu

  subroutine build_case(d)
    type(hugks_desc), intent(out) :: d
    real(8), allocatable :: roots(:), wgh(:)
    real(8) :: umax_dt, du, innerk
    integer :: i

    dx = xlength / nx                           ! 1d/src/initialization.f90:68
    allocate(uspace(nv), weight(nv), roots(nv), wgh(nv))
    if (trim(quad_type) == 'gauss_hermite') then
      call gauss_hermite(nv, Tref, PI, uspace, weight, roots, wgh)
      umax_dt = uspace(nv)
      du = (uspace(nv) - uspace(1)) / (nv - 1)    ! 1d/src/velocity.f90:146
      d%fd_span = uspace(nv) - uspace(1)          ! lengthU, 1d/src/velocity.f90:147
    else if (trim(quad_type) == 'newton_cotes') then
      du = (2.d0 * vmax * sqrt(Tref)) / (nv - 1)
      do i = 1, nv
        uspace(i) = -vmax * sqrt(Tref) + (i - 1) * du
        weight(i) = newton_cotes_coeff(i, nv) * du
      end do
      umax_dt = vmax * sqrt(Tref)
      d%fd_span = vmax * sqrt(Tref) - (-vmax * sqrt(Tref))   ! lengthU, 1d/src/config.f90:75-78
    else
      write(*, *) "unknown quad_type '", trim(quad_type), "'"
      error stop 1
    end if

    innerk = 2.d0
    gam = (innerk + 3.d0) / (innerk + 1.d0)      ! 1d/src/gastheory.f90:60-66
    d%abi_version = HUGKS_ABI_VERSION
    d%dim = 1
    d%nx = nx
    d%ny = 1
    d%nv = nv
    d%nm = Nm
    select case (trim(precision))
    case ('fp64')
      d%precision = HUGKS_FP64
    case ('mixed', 'mixed_fp32')
      d%precision = HUGKS_MIXED_FP32
    case ('mixed_bf16x3', 'bf16x3')
      d%precision = HUGKS_MIXED_BF16X3
    case default
      write(*, *) "unknown precision '", trim(precision), "'"
      error stop 1
    end select
    select case (trim(force_type))              ! 1d/src/solver.f90:233-242
    case ('hermite')
      d%force_variant = HUGKS_FORCE_HERMITE_LEGACY
    case ('finite_difference')
      d%force_variant = HUGKS_FORCE_FINITE_DIFFERENCE
    case default
      write(*, *) "unknown force_type '", trim(force_type), "'"
      error stop 1
    end select
    d%clamp_dt = 0
    d%dx = dx
    d%dy = 1.d0
    d%pi = PI
    d%smv = tiny(1.d0)
    d%innerk = innerk
    d%gamma = gam
    d%prandtl = PRANDTL
    d%shakhov_coef = dble(0.8)
    d%omega = 0.72d0
    d%muref = get_mu(knudsen, 1.d0, 0.5d0)
    d%tau_model = HUGKS_TAU_VHS
    d%dt_rule = HUGKS_DT_1D
    d%cfl = cfl
    d%maxtime = maxtime
    d%eps_res = 1.d-6
    d%umax_dt = umax_dt
    d%vmax_dt = umax_dt
    d%force_rate = abs(phi) / du
    d%tref = Tref
    d%phi = phi
    d%tw = [TwL, TwR, 1.d0, 1.d0]

    allocate(xc(nx), prim0(3, nx))
    prim0 = 0.d0
    do i = 1, nx
      xc(i) = (i - 0.5d0) * dx
    end do
    select case (case_id)
    case (2)   ! hydrostatic equilibrium with a pressure bump
      do i = 1, nx
        prim0(1, i) = exp(-xc(i))
        prim0(3, i) = 0.5d0 * prim0(1, i) / (exp(-xc(i)) + 0.01d0 * exp(-100.d0 * (xc(i) - 0.5d0)**2))
      end do
      d%bc = [HUGKS_BC_SPECULAR, HUGKS_BC_SPECULAR, 0, 0]
    case (3)   ! Sod shock tube
      do i = 1, nx
        if (i <= nx / 2) then
          prim0(:, i) = [1.d0, 0.d0, 0.5d0]
        else
          prim0(:, i) = [0.125d0, 0.d0, 0.625d0]
        end if
      end do
      d%bc = [HUGKS_BC_SPECULAR, HUGKS_BC_SPECULAR, 0, 0]
    case (4)   ! Fourier flow between diffuse walls
      do i = 1, nx
        prim0(1, i) = 1.d0
        prim0(3, i) = 1.d0 / (TwL + xc(i) * (TwR - TwL))
      end do
      d%bc = [HUGKS_BC_DIFFUSE, HUGKS_BC_DIFFUSE, 0, 0]
      d%steady_steps = steady_steps
      d%steady_tol = steady_tol
      d%mass_tol = mass_tol
    case default
      write(*, *) "Error: unknown case_id ", case_id
      error stop 1
    end select
    d%u = c_loc(uspace)
    d%weight = c_loc(weight)
    d%prim0 = c_loc(prim0)
  end subroutine build_case

  ! Reference pressure of the hydrostatic case read as the reference's
  ! output does (1d/src/output.f90:17-53): 270 header lines, then 600
  ! values; absent file: exp(-x).
  subroutine reference_pressure(pref, have)
    real(8), intent(out) :: pref(:)
    logical, intent(out) :: have
    integer :: u, i, j, k, ios
    logical :: found
    real(8) :: buf(600)
    have = .false.
    if (len_trim(refpfile) == 0) return
    inquire(file=refpfile, exist=found)
    if (.not. found) then
      write(*, *) "reference file '", trim(refpfile), "' not found; using exp(-x)"
      return
    end if
    open(newunit=u, file=refpfile, status='old', action='read')
    do i = 1, 270
      read(u, *)
    end do
    k = 0
    do i = 1, 66
      read(u, *, iostat=ios) (buf(k + j), j = 1, 9)
      if (ios /= 0) then
        close(u)
        return
      end if
      k = k + 9
    end do
    read(u, *, iostat=ios) (buf(594 + j), j = 1, 6)
    close(u)
    pref = buf(1:size(pref))
    have = .true.
  end subroutine reference_pressure

  ! write_macro of 1d/src/output.f90:55-103.
  subroutine write_macro(h, loop, stepped)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: loop
    logical, intent(in) :: stepped
    real(8), allocatable, target :: w(:, :), q(:)
    real(8), allocatable :: sol(:, :), pref(:)
    real(c_double), allocatable, target :: vals(:)
    integer(c_int64_t), target :: lens(10)
    real(8) :: rho, uu, lam
    character(len=8) :: ctemp
    character(len=80) :: zone
    logical :: have
    integer :: i

    allocate(w(3, nx), q(nx), sol(9, nx), pref(nx))
    call hugks_check(hugks_get_macro(h, c_loc(w), c_loc(q)), h, 'get_macro')
    have = .false.
    if (case_id == 2) call reference_pressure(pref, have)
    do i = 1, nx
      rho = w(1, i)
      uu = w(2, i) / w(1, i)
      lam = 0.5d0 * w(1, i) / (gam - 1.d0) / (w(3, i) - 0.5d0 * w(2, i)**2 / w(1, i))
      sol(1, i) = rho
      sol(2, i) = uu
      sol(3, i) = 1.d0 / lam
      sol(4, i) = RC * sol(1, i) * sol(3, i)
      if (.not. (case_id == 2 .and. have)) pref(i) = exp(-xc(i))
      sol(5, i) = sol(4, i) - pref(i)
      sol(6, i) = q(i)
      ! error1, error2, N_order1 of the force module: 0 before the first
      ! step, then 1, 1, Nm (1d/src/gravity.f90:11-13,68-69; solver.f90:237)
      if (stepped) then
        sol(7:8, i) = 1.d0
        sol(9, i) = Nm
      else
        sol(7:9, i) = 0.d0
      end if
    end do
    ! the file of output.f90:89-99 -- the list-directed VARIABLES line (a
    ! leading blank), the ZONE line, then the blocks in the
    ! "(9(ES23.16,2X))" form -- is written by the library's writer thread
    write(ctemp, '(i8)') loop
    write(zone, '(A,I5,A,I5,A)') 'ZONE I =', nx, ', DATAPACKING=BLOCK'
    allocate(vals(10 * nx))
    vals(1:nx) = xc
    do i = 1, 9
      vals(i * nx + 1:(i + 1) * nx) = sol(i, :)
    end do
    lens = nx
    call hugks_check(hugks_write_blocks_async(h, &
        'macro'//trim(adjustl(ctemp))//'.dat'//c_null_char, &
        ' VARIABLES = X, RHO, U, T, P, PERTURBATION, Q, error1, error2, N' // &
        new_line('a')//trim(zone)//new_line('a')//c_null_char, &
        9_c_int32_t, 10_c_int32_t, c_loc(lens), c_loc(vals)), h, 'write_macro')
  end subroutine write_macro

end module ugks1d_setup

program ugks1d_gpu
  use, intrinsic :: iso_c_binding
  use hugks_capi
  use ugks1d_setup
  implicit none
  character(len=256) :: cfgfile
  type(hugks_desc) :: desc
  type(c_ptr) :: h
  type(hugks_step_info) :: info
  integer :: iter
  integer(8) :: c0, c1, rate

  cfgfile = 'input.namelist'
  if (command_argument_count() >= 1) call get_command_argument(1, cfgfile)
  call read_config(cfgfile)
  write(*, *) "UGKS + Hermite 1D solver (libhugks), case_id =", case_id
  write(*, *) "device: ", hugks_device_info()
  call system_clock(c0, rate)
  call build_case(desc)
  call hugks_check(hugks_create(desc, h), c_null_ptr, 'create')
  iter = 1
  call write_macro(h, iter, .false.)

  ! main loop (1d/src/main.f90:79-127): maxtime and the Fourier steady
  ! test are evaluated on the device
  write(*, *) "begin main loop"
  call hugks_check(hugks_advance(h, int(max_steps, c_int32_t), info), h, 'advance')
  iter = info%iter
  if (info%stop_reason == HUGKS_STOP_STEADY) then
    write(*, '(A,I8,3ES15.7)') 'Fourier steady state reached: ', iter, info%simtime, &
        maxval(abs(info%steady_res)), abs(info%mass_rel_change)
  end if
  write(*, *) 'end of mainLoop'
  write(*, *) 'iteration: ', iter
  if (case_id == 4 .and. info%stop_reason /= HUGKS_STOP_STEADY) then
    write(*, *) 'WARNING: Fourier case reached maxtime before steady tolerance'
  end if
  call write_macro(h, iter, info%steps > 0)
  call hugks_check(hugks_write_wait(h), h, 'write_macro')
  call system_clock(c1)
  write(*, *) "Elapsed time: ", real(c1 - c0) / real(rate), " seconds"
  call hugks_check(hugks_destroy(h), h, 'destroy')
end program ugks1d_gpu
