! ------------------------------------------------------------
! 2D Hermite-UGKS driver over libhugks.so.
!
! Usage: ugks2d_gpu [input.namelist]
!
! The input is the reference's &params namelist (2d/src/config.f90:47-48,
! same names and defaults) optionally followed by a &gpu group
! (precision = 'fp64' | 'mixed' | 'bf16x3'; max_steps, 0 = no cap; rt_variant =
! 'reference' | 'cartesian'; seed; restart_file, checkpoint_file: continue
! from / save the whole solver state, hugks_checkpoint_read/_write).  The driver keeps what the reference
! main program does around the step - case selection, velocity
! quadrature, initial condition, body force, the loop exits and output
! times, macro<iter>.dat files and the final mass report
! (2d/src/main.f90:13-137) - and hands the step itself to the library.
! ------------------------------------------------------------
module ugks2d_setup
  use, intrinsic :: iso_c_binding
  use hugks_capi
  use hugks_quadrature
  implicit none

  ! &params, defaults of 2d/src/config.f90:11-40
  integer :: nx = 3, ny = 100
  real(8) :: xlength = 0.1d0, ylength = 1.0d0
  integer :: nv = 16
  real(8) :: vmax = 5.0d0, Tref = 1.0d0
  character(len=20) :: quad_type = 'gauss_hermite'
  character(len=20) :: force_type = 'hermite_symmetric'
  integer :: Nm = 7
  real(8) :: cfl = 0.8d0, maxtime = 200.d0
  integer :: case_id = 5
  real(8) :: knudsen = 2.d-2, phi = 0.1d0, TwL = 1.0d0, TwR = 1.0d0
  namelist /params/ nx, ny, xlength, ylength, nv, vmax, Tref, quad_type, &
                    force_type, Nm, cfl, maxtime, case_id, knudsen, phi, TwL, TwR

  ! &gpu
  character(len=16) :: precision = 'fp64'
  integer :: max_steps = 0
  character(len=16) :: rt_variant = 'reference'
  integer :: seed = 0
  character(len=256) :: restart_file = '', checkpoint_file = ''
  namelist /gpu/ precision, max_steps, rt_variant, seed, restart_file, checkpoint_file

  ! constants of the reference (2d/src/data2d.f90:9-11,54)
  real(8), parameter :: PI = 3.1415926535d0
  real(8), parameter :: SMV = 1.d-8
  real(8), parameter :: EPS = 1.d-7
  real(8), parameter :: RC = 0.5d0

  ! arrays handed to the library (kept alive for c_loc)
  real(8), allocatable, target :: uspace(:), vspace(:), weight(:, :)
  real(8), allocatable, target :: prim0(:, :, :)          ! (4, nx, ny)
  real(8), allocatable, target :: acc_cell(:, :, :)       ! (2, nx, ny)
  real(8), allocatable, target :: acc_v(:, :, :)          ! (2, nx+1, ny)
  real(8), allocatable, target :: acc_h(:, :, :)          ! (2, nx, ny+1)
  real(8), allocatable :: xc(:), yc(:)
  real(8) :: dx, dy, time2, time3, time4

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
    read(u, nml=gpu, iostat=ios)        ! optional group
    close(u)
  end subroutine read_config

  ! Hard-sphere/VHS reference viscosity (2d/src/gastheory2d.f90:101-111).
  real(8) function get_mu(kn, alpha, omega)
    real(8), intent(in) :: kn, alpha, omega
    get_mu = 5.d0 * (alpha + 1.d0) * (alpha + 2.d0) * sqrt(PI) / &
             (4.d0 * alpha * (5.d0 - 2.d0 * omega) * (7.d0 - 2.d0 * omega)) * kn
  end function get_mu

  ! Uniform (0,1) value of cell (i, j) for the Cartesian RT noise:
  ! Park-Miller MINSTD, three steps from a hashed start.
  real(8) function cell_uniform(s, i, j)
    integer, intent(in) :: s, i, j
    integer(8) :: x
    integer :: k
    x = mod(int(s, 8) * 1000003_8 + int(j, 8) * 7919_8 + int(i, 8) * 104729_8 + 12345_8, &
            2147483646_8) + 1_8
    do k = 1, 3
      x = mod(48271_8 * x, 2147483647_8)
    end do
    cell_uniform = dble(x) / 2147483647.d0
  end function cell_uniform

  subroutine build_case(d)
    type(hugks_desc), intent(out) :: d
    real(8), allocatable :: roots(:), wgh(:), w1(:), ctheta(:, :), vtheta(:, :), htheta(:, :)
    real(8) :: umax_dt, vmax_dt, gap, du, r, th, radius, g, yi, innerk
    integer :: i, j

    dx = xlength / nx                         ! 2d/src/initialization.f90:69-70
    dy = ylength / ny
    allocate(uspace(nv), vspace(nv), weight(nv, nv), roots(nv), wgh(nv), w1(nv))
    if (trim(quad_type) == 'gauss_hermite') then
      call gauss_hermite(nv, Tref, PI, uspace, w1, roots, wgh)
      vspace = uspace
      do j = 1, nv
        do i = 1, nv
          weight(i, j) = wgh(i) * exp(roots(i)**2) * sqrt(Tref) * wgh(j) * exp(roots(j)**2) &
                         * sqrt(Tref)
        end do
      end do
      umax_dt = uspace(nv)
      vmax_dt = vspace(nv)
      d%fd_span = uspace(nv) - uspace(1)      ! lengthU, 2d/src/velocity.f90:131-134
    else if (trim(quad_type) == 'newton_cotes') then
      du = (2.d0 * vmax * sqrt(Tref)) / (nv - 1)
      do i = 1, nv
        uspace(i) = -vmax * sqrt(Tref) + (i - 1) * du
      end do
      vspace = uspace
      do j = 1, nv
        do i = 1, nv
          weight(i, j) = (newton_cotes_coeff(i, nv) * du) * (newton_cotes_coeff(j, nv) * du)
        end do
      end do
      umax_dt = vmax * sqrt(Tref)
      vmax_dt = umax_dt
      d%fd_span = vmax * sqrt(Tref) - (-vmax * sqrt(Tref))   ! lengthU, 2d/src/config.f90:75-78
    else
      write(*, *) "unknown quad_type '", trim(quad_type), "'"
      error stop 1
    end if
    gap = huge(1.d0)                          ! 2d/src/solver.f90:53-54
    do i = 1, nv - 1
      gap = min(gap, uspace(i + 1) - uspace(i), vspace(i + 1) - vspace(i))
    end do

    innerk = 1.d0
    d%abi_version = HUGKS_ABI_VERSION
    d%dim = 2
    d%nx = nx
    d%ny = ny
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
    select case (trim(force_type))
    case ('hermite_symmetric')
      d%force_variant = HUGKS_FORCE_HERMITE_SYMMETRIC
    case ('hermite_symmetric_b')
      d%force_variant = HUGKS_FORCE_HERMITE_SYMMETRIC_B
    case ('hermite')
      d%force_variant = HUGKS_FORCE_HERMITE_LEGACY
    case ('finite_difference')
      d%force_variant = HUGKS_FORCE_FINITE_DIFFERENCE
    case default
      write(*, *) "unknown force_type '", trim(force_type), "'"
      error stop 1
    end select
    d%clamp_dt = 1
    d%dx = dx
    d%dy = dy
    d%pi = PI
    d%smv = SMV
    d%innerk = innerk
    d%gamma = (innerk + 4.d0) / (innerk + 2.d0)   ! 2d/src/gastheory2d.f90:70-76
    d%shakhov_coef = dble(0.8)                    ! default-REAL literal of gastheory2d.f90:22
    d%omega = 0.72d0
    d%cfl = cfl
    d%maxtime = maxtime
    d%eps_res = EPS
    d%umax_dt = umax_dt
    d%vmax_dt = vmax_dt
    d%force_rate = abs(phi) / gap
    d%tref = Tref

    allocate(xc(nx), yc(ny))
    do i = 1, nx
      xc(i) = (i - 0.5d0) * dx
    end do
    do j = 1, ny
      yc(j) = (j - 0.5d0) * dy
    end do
    allocate(prim0(4, nx, ny), acc_cell(2, nx, ny), acc_v(2, nx + 1, ny), acc_h(2, nx, ny + 1))
    prim0 = 0.d0
    allocate(ctheta(nx, ny), vtheta(nx + 1, ny), htheta(nx, ny + 1))

    select case (case_id)
    case (5)
      ! x-periodic, diffuse walls in y (2d/src/solver.f90:204-236)
      d%bc = [HUGKS_BC_PERIODIC, HUGKS_BC_PERIODIC, HUGKS_BC_DIFFUSE, HUGKS_BC_DIFFUSE]
      d%tw = [1.d0, 1.d0, TwL, TwR]
      d%tau_model = HUGKS_TAU_MU_OVER_RHO
      d%dt_rule = HUGKS_DT_2D_WALL_NORMAL
      d%stop_on_residual = 1
      d%prandtl = 1.d0
      d%muref = 0.5d0 * sqrt(PI) * knudsen
      prim0(1, :, :) = 1.d0
      prim0(4, :, :) = 1.d0 / Tref
      ! uniform force phi along x (angle 0)
      ctheta = 0.d0
      vtheta = 0.d0
      htheta = 0.d0
    case (6)
      d%bc = HUGKS_BC_SPECULAR
      d%tw = 1.d0
      d%tau_model = HUGKS_TAU_VHS
      d%dt_rule = HUGKS_DT_2D_GENERAL
      d%stop_on_residual = 0
      d%prandtl = 2.d0 / 3.d0
      d%muref = get_mu(knudsen, 1.d0, 0.5d0)
      if (trim(rt_variant) == 'reference') then
        ! polar initial condition, 2d/src/initialization.f90:146-170
        do j = 1, ny
          do i = 1, nx
            r = sqrt(xc(i)**2 + yc(j)**2)
            th = atan2(yc(j), xc(i))
            radius = 0.6d0 * (1.d0 + 0.02d0 * cos(20 * th))
            if (r <= radius) then
              prim0(1, i, j) = exp(-2.68d0 * (r + 0.258d0))
            else
              prim0(1, i, j) = exp(-5.53d0 * (r - 0.308d0))
            end if
            if (r <= 0.62324965d0) then
              prim0(4, i, j) = 0.5d0 * prim0(1, i, j) / 1.5d0 / exp(-2.68d0 * (r + 0.258d0)) * 2.68d0
            else
              prim0(4, i, j) = 0.5d0 * prim0(1, i, j) / 1.5d0 / exp(-5.53d0 * (r - 0.308d0)) * 5.53d0
            end if
            ctheta(i, j) = th
            vtheta(i, j) = atan2(yc(j), xc(i) - 0.5d0 * dx)       ! initialization.f90:86-114
            htheta(i, j) = atan2(yc(j) - 0.5d0 * dy, xc(i))
          end do
          vtheta(nx + 1, j) = atan2(yc(j), xc(nx) + 0.5d0 * dx)
        end do
        do i = 1, nx
          htheta(i, ny + 1) = atan2(yc(ny) + 0.5d0 * dy, xc(i))
        end do
      else if (trim(rt_variant) == 'cartesian') then
        ! heavy isothermal layer (T=1) over a light one (T=2), density ratio
        ! 2 at y_i(x) = ylength/2 + 0.01 cos(2 pi x / xlength), gravity
        ! (0, -|phi|), y-velocity noise of +-1e-4 per cell
        g = abs(phi)
        phi = g
        do j = 1, ny
          do i = 1, nx
            yi = 0.5d0 * ylength + 0.01d0 * cos(2.d0 * acos(-1.d0) * xc(i) / xlength)
            if (yc(j) >= yi) then
              prim0(1, i, j) = exp(-2.d0 * g * (yc(j) - yi) / 1.d0)
              prim0(4, i, j) = 1.d0
            else
              prim0(1, i, j) = 0.5d0 * exp(2.d0 * g * (yi - yc(j)) / 2.d0)
              prim0(4, i, j) = 0.5d0
            end if
            prim0(3, i, j) = 1.d-4 * (2.d0 * cell_uniform(seed, i, j) - 1.d0)
          end do
        end do
        ctheta = atan2(-1.d0, 0.d0)
        vtheta = ctheta(1, 1)
        htheta = ctheta(1, 1)
      else
        write(*, *) "unknown rt_variant '", trim(rt_variant), "'"
        error stop 1
      end if
    case default
      write(*, *) "Error: unknown case_id ", case_id
      error stop 1
    end select
    d%phi = phi

    ! cell force (2d/src/solver.f90:291-297) and face force projected on
    ! the face normal/tangent (2d/src/flux.f90:12-26): vertical faces
    ! (cosa, sina) = (1, 0), horizontal (0, 1)
    acc_cell(1, :, :) = phi * cos(ctheta)
    acc_cell(2, :, :) = phi * sin(ctheta)
    acc_v(1, :, :) = phi * cos(vtheta) * 1.d0 + phi * sin(vtheta) * 0.d0
    acc_v(2, :, :) = phi * sin(vtheta) * 1.d0 - phi * cos(vtheta) * 0.d0
    acc_h(1, :, :) = phi * cos(htheta) * 0.d0 + phi * sin(htheta) * 1.d0
    acc_h(2, :, :) = phi * sin(htheta) * 0.d0 - phi * cos(htheta) * 1.d0

    d%u = c_loc(uspace)
    d%v = c_loc(vspace)
    d%weight = c_loc(weight)
    d%prim0 = c_loc(prim0)
    d%accel_cell = c_loc(acc_cell)
    d%accel_vface = c_loc(acc_v)
    d%accel_hface = c_loc(acc_h)

    ! output times of the RT case (2d/src/initialization.f90:52-60)
    if (case_id == 6 .and. maxtime <= 0.2d0) then
      time2 = 0.04d0
      time3 = 0.08d0
      time4 = 0.14d0
    else
      time2 = 0.8d0
      time3 = 1.4d0
      time4 = 2.d0
    end if
  end subroutine build_case

  ! write_macro of 2d/src/output2d.f90:8-56: Tecplot block file
  ! macro<loop>.dat with X, Y at grid nodes and the cell-centred fields.
  subroutine write_macro(h, loop)
    type(c_ptr), intent(in) :: h
    integer, intent(in) :: loop
    real(8), allocatable, target :: w(:, :, :), q(:, :, :)
    real(8), allocatable :: sol(:, :, :), coord(:, :, :)
    real(c_double), allocatable, target :: vals(:)
    integer(c_int64_t), target :: lens(13)
    real(8) :: prim(4)
    character(len=11) :: ctemp
    character(len=160) :: zone
    integer :: i, j, np1, nc

    allocate(w(4, nx, ny), q(2, nx, ny), sol(11, nx, ny), coord(2, nx + 1, ny + 1))
    call hugks_check(hugks_get_macro(h, c_loc(w), c_loc(q)), h, 'get_macro')
    do j = 1, ny
      do i = 1, nx
        call primitive(w(:, i, j), prim)
        sol(1:3, i, j) = prim(1:3)
        sol(4, i, j) = 1.d0 / prim(4)
        sol(5, i, j) = 0.5 * prim(1) / prim(4)
        sol(6:7, i, j) = q(:, i, j)
        sol(8, i, j) = sqrt(xc(i)**2 + yc(j)**2)
        sol(11, i, j) = 1
      end do
    end do
    do j = 1, ny
      do i = 1, nx
        call primitive(w(:, i, 1), prim)
        sol(9, i, j) = sqrt(xc(i)**2 + yc(1)**2)
        sol(10, i, j) = prim(1)
      end do
    end do
    do j = 1, ny + 1
      do i = 1, nx + 1
        coord(1, i, j) = (i - 1) * dx
        coord(2, i, j) = (j - 1) * dy
      end do
    end do
    ! the file of output2d.f90:38-55 -- the list-directed VARIABLES line
    ! (a leading blank), the ZONE line, then the blocks in the
    ! "(11(ES23.16,2X))" form -- is written by the library's writer thread
    ! while the run goes on
    write(ctemp, '(i11)') loop
    write(zone, '(A,I5,A,I5,A)') 'ZONE I =', nx + 1, ', J =', ny + 1, &
        ', DATAPACKING=BLOCK,VARLOCATION=([3-13]=CELLCENTERED)'
    np1 = (nx + 1) * (ny + 1)
    nc = nx * ny
    allocate(vals(2 * np1 + 11 * nc))
    vals(1:np1) = reshape(coord(1, :, :), [np1])
    vals(np1 + 1:2 * np1) = reshape(coord(2, :, :), [np1])
    do i = 1, 11
      vals(2 * np1 + (i - 1) * nc + 1:2 * np1 + i * nc) = reshape(sol(i, :, :), [nc])
    end do
    lens(1:2) = np1
    lens(3:13) = nc
    call hugks_check(hugks_write_blocks_async(h, &
        'macro'//trim(adjustl(ctemp))//'.dat'//c_null_char, &
        ' VARIABLES = X, Y, RHO, U, V, T, P, QX, QY, radius, radius-0, RHO-0,error' // &
        new_line('a')//trim(zone)//new_line('a')//c_null_char, &
        11_c_int32_t, 13_c_int32_t, c_loc(lens), c_loc(vals)), h, 'write_macro')
  end subroutine write_macro

  ! get_primary, 2d/src/gastheory2d.f90:28-36 (gamma of K = 1)
  subroutine primitive(w, prim)
    real(8), intent(in) :: w(4)
    real(8), intent(out) :: prim(4)
    real(8) :: gam
    gam = 5.d0 / 3.d0
    prim(1) = w(1)
    prim(2) = w(2) / w(1)
    prim(3) = w(3) / w(1)
    prim(4) = 0.5d0 * w(1) / (gam - 1.d0) / (w(4) - 0.5d0 * (w(2)**2 + w(3)**2) / w(1))
  end subroutine primitive

end module ugks2d_setup

program ugks2d_gpu
  use, intrinsic :: iso_c_binding
  use hugks_capi
  use ugks2d_setup
  implicit none
  character(len=256) :: cfgfile
  type(hugks_desc) :: desc
  type(c_ptr) :: h
  type(hugks_step_info) :: info
  real(8), allocatable, target :: w(:, :, :)
  real(8) :: mass_initial, mass_final, max_u_axis, max_v_axis, prim(4), t_next
  integer :: iter, i, j, k, taken, cap, chunk, stop_reason
  real(c_double) :: t_restart = 0.d0
  integer(c_int32_t) :: it_restart
  logical :: wrote(3)
  integer(8) :: c0, c1, rate

  cfgfile = 'input.namelist'
  if (command_argument_count() >= 1) call get_command_argument(1, cfgfile)
  call read_config(cfgfile)
  write(*, *) "UGKS + Hermite 2D solver (libhugks), case_id =", case_id
  write(*, *) "device: ", hugks_device_info()
  call system_clock(c0, rate)

  call build_case(desc)
  call hugks_check(hugks_create(desc, h), c_null_ptr, 'create')
  iter = 1
  if (len_trim(restart_file) > 0) then
    ! continue a saved run: its iteration index and time; the initial mass
    ! below is then the mass at the restart point
    call hugks_check(hugks_checkpoint_read(h, restart_file), h, 'restart')
    call hugks_check(hugks_get_time(h, t_restart, it_restart), h, 'get_time')
    iter = it_restart
    write(*, *) "restarted from ", trim(restart_file), " at iteration ", iter, " time ", t_restart
  else
    call write_macro(h, iter)
  end if
  allocate(w(4, nx, ny))
  call hugks_check(hugks_get_macro(h, c_loc(w), c_null_ptr), h, 'get_macro')
  mass_initial = 0.d0
  do j = 1, ny
    do i = 1, nx
      mass_initial = mass_initial + w(1, i, j) * (dx * dy)
    end do
  end do

  ! main loop (2d/src/main.f90:76-108): each library call runs until the
  ! next output point, the step cap or a loop exit
  write(*, *) "begin main loop"
  wrote = .false.
  if (len_trim(restart_file) > 0 .and. case_id == 6) then
    ! output times already passed in the saved run
    wrote(1) = t_restart >= time2
    wrote(2) = t_restart >= time3
    wrote(3) = t_restart >= time4
  end if
  taken = 0
  stop_reason = HUGKS_STOP_NONE
  do
    cap = 0
    if (max_steps > 0) cap = max_steps - taken
    if (max_steps > 0 .and. cap <= 0) exit
    if (case_id == 6) then
      t_next = huge(1.d0)
      if (.not. wrote(1)) then
        t_next = time2
      else if (.not. wrote(2)) then
        t_next = time3
      else if (.not. wrote(3)) then
        t_next = time4
      end if
      call hugks_check(hugks_advance_until(h, int(cap, c_int32_t), t_next, info), h, 'advance')
    else
      ! the next iteration index that is a multiple of 5000 writes a file
      chunk = mod(5000 - mod(iter, 5000), 5000) + 1
      if (cap > 0) chunk = min(chunk, cap)
      call hugks_check(hugks_advance(h, int(chunk, c_int32_t), info), h, 'advance')
    end if
    taken = taken + info%steps
    if (info%steps == 0) exit
    ! the last step's iteration index before the reference's iter=iter+1
    k = info%iter - 1
    if (case_id == 6) then
      if (.not. wrote(1) .and. info%simtime >= time2) then
        call write_macro(h, k)
        wrote(1) = .true.
      end if
      if (.not. wrote(2) .and. info%simtime >= time3) then
        call write_macro(h, k)
        wrote(2) = .true.
      end if
      if (.not. wrote(3) .and. info%simtime >= time4) then
        call write_macro(h, k)
        wrote(3) = .true.
      end if
    else if (mod(k, 5000) == 0) then
      call write_macro(h, k)
    end if
    iter = info%iter
    write(*, "(A24,I8,6E15.7)") "iteration,simtime,dt,res:", k, info%simtime, info%dt, info%res
    if (info%stop_reason == HUGKS_STOP_MAXTIME .or. info%stop_reason == HUGKS_STOP_RESIDUAL) then
      stop_reason = info%stop_reason
      exit
    end if
  end do

  write(*, *) 'end of mainLoop'
  write(*, *) 'iteration: ', iter
  write(*, *) "equi time vs Knudsen: ", info%simtime, knudsen
  call write_macro(h, iter)
  call hugks_check(hugks_get_macro(h, c_loc(w), c_null_ptr), h, 'get_macro')
  mass_final = 0.d0
  max_u_axis = 0.d0
  max_v_axis = 0.d0
  do j = 1, ny
    do i = 1, nx
      mass_final = mass_final + w(1, i, j) * (dx * dy)
    end do
    call primitive(w(:, 1, j), prim)
    max_u_axis = max(max_u_axis, abs(prim(2)))
  end do
  do i = 1, nx
    call primitive(w(:, i, 1), prim)
    max_v_axis = max(max_v_axis, abs(prim(3)))
  end do
  write(*, '(A,ES24.16)') 'initial_mass = ', mass_initial
  write(*, '(A,ES24.16)') 'final_mass   = ', mass_final
  write(*, '(A,ES24.16)') 'relative_mass_change = ', (mass_final - mass_initial) / mass_initial
  write(*, '(A,ES24.16)') 'max_abs_u_left_axis  = ', max_u_axis
  write(*, '(A,ES24.16)') 'max_abs_v_bottom_axis= ', max_v_axis
  call system_clock(c1)
  write(*, *) "Elapsed time: ", real(c1 - c0) / real(rate), " seconds"
  if (len_trim(checkpoint_file) > 0) then
    call hugks_check(hugks_checkpoint_write(h, checkpoint_file), h, 'checkpoint')
    write(*, *) "state saved to ", trim(checkpoint_file)
  end if
  call hugks_check(hugks_write_wait(h), h, 'write_macro')
  call hugks_check(hugks_destroy(h), h, 'destroy')
end program ugks2d_gpu
