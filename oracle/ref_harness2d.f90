! ------------------------------------------------------------------
! ref_harness2d: test-infrastructure driver around the reference 2D
! solver modules (compiled unmodified from the reference sources by
! oracle/Makefile).  It is the FP64 oracle and the CPU baseline; nothing
! in the product links or calls it.
!
! It replaces only the reference main program (2d/src/main.f90:13-137):
!   * same phase sequence per step: timestep, interpolation, evolution,
!     update, simtime += dt, iter += 1 (2d/src/main.f90:78-82,102)
!   * same exits: simtime >= maxtime, case 5 all(res < eps)
!     (2d/src/main.f90:103-104), plus a step cap
!   * optional Cartesian Rayleigh-Taylor initial condition, injected
!     after the reference initialize() by overwriting prim/w/h/b and the
!     cell/face angles the reference uses for the acceleration direction
!     (2d/src/flux.f90:21-22, 2d/src/solver.f90:292-293)
!   * per-phase wall time, and a binary dump of the final state.
!
! usage: ref_harness2d <namelist> <dump|-> <max_steps> <rt_cartesian 0|1>
!                      <seed> [sub_ny]
!   sub_ny > 0 keeps only the lowest sub_ny rows of the domain for
!   timing samples (the run is then a bounded sample, not a parity run).
!
! dump layout (stream, native endian):
!   int32 nx, ny, nv, steps, iter
!   real64 simtime, dt, res(4)
!   real64 dts(steps)
!   real64 u(nv), v(nv), weight(nv*nv)        (Fortran order)
!   real64 w(4, nx, ny)
!   real64 h(nv*nv, nx, ny), b(nv*nv, nx, ny)
! ------------------------------------------------------------------
program ref_harness2d
use CONFIG
use ALLDATA
use GRIDMAKER
use INITIALIZER
use SOLVER
implicit none

character(len=256) :: cfgfile, dumpfile, arg
integer :: max_steps, rt_cart, seed, steps, i, j, u, sub_ny
integer(8) :: c0, c1, crate, tphase(4), tloop0, tloop1
real(8), allocatable :: dts(:)
real(8) :: mass0, mass1

call get_command_argument(1, cfgfile)
call get_command_argument(2, dumpfile)
call get_command_argument(3, arg); read(arg,*) max_steps
call get_command_argument(4, arg); read(arg,*) rt_cart
call get_command_argument(5, arg); read(arg,*) seed
sub_ny = 0
if (command_argument_count() >= 6) then
  call get_command_argument(6, arg); read(arg,*) sub_ny
end if

call read_config(trim(cfgfile))
isRestart = .false.
iniType = 1
gridfile = 'none'
if (sub_ny > 0 .and. sub_ny < ny) then
  ! keep the cell size of the full grid
  ylength = ylength*dble(sub_ny)/dble(ny)
  ny = sub_ny
end if
call gridInit()
call initialize()
if (rt_cart /= 0) call init_rt_cartesian(seed)

allocate(dts(max(max_steps,1)))
mass0 = 0.d0
do j=iymin,iymax
  do i=ixmin,ixmax
    mass0 = mass0 + ctr(i,j)%w(1)*ctr(i,j)%area
  end do
end do

call system_clock(count_rate=crate)
tphase = 0
steps = 0
call system_clock(tloop0)
do while (steps < max_steps)
  call system_clock(c0); call timestep();      call system_clock(c1); tphase(1)=tphase(1)+c1-c0
  call system_clock(c0); call interpolation(); call system_clock(c1); tphase(2)=tphase(2)+c1-c0
  call system_clock(c0); call evolution();     call system_clock(c1); tphase(3)=tphase(3)+c1-c0
  call system_clock(c0); call update();        call system_clock(c1); tphase(4)=tphase(4)+c1-c0
  simtime = simtime + dt
  steps = steps + 1
  dts(steps) = dt
  iter = iter + 1
  if (simtime >= maxtime) exit
  if (case_id == 5 .and. all(res < eps)) exit
end do
call system_clock(tloop1)

mass1 = 0.d0
do j=iymin,iymax
  do i=ixmin,ixmax
    mass1 = mass1 + ctr(i,j)%w(1)*ctr(i,j)%area
  end do
end do

write(*,'(A,I8,A,I8,A,ES24.16,A,ES24.16)') 'HARNESS steps=', steps, ' iter=', iter, &
     ' simtime=', simtime, ' rel_mass_change=', (mass1-mass0)/mass0
write(*,'(A,ES14.6,A,4ES14.6)') 'HARNESS_TIMING loop_s=', dble(tloop1-tloop0)/dble(crate), &
     ' phases_s=', dble(tphase)/dble(crate)

if (trim(dumpfile) /= '-') then
  open(newunit=u, file=trim(dumpfile), access='stream', form='unformatted', status='replace')
  write(u) int(nx,4), int(ny,4), int(nv,4), int(steps,4), int(iter,4)
  write(u) simtime, dt, res
  write(u) dts(1:steps)
  write(u) uspace(:,1), vspace(1,:), weight
  do j=iymin,iymax
    do i=ixmin,ixmax
      write(u) ctr(i,j)%w
    end do
  end do
  do j=iymin,iymax
    do i=ixmin,ixmax
      write(u) ctr(i,j)%h
    end do
  end do
  do j=iymin,iymax
    do i=ixmin,ixmax
      write(u) ctr(i,j)%b
    end do
  end do
  close(u)
end if

contains

! Park-Miller MINSTD step; all values stay below 2**47 in int64.
integer(8) function minstd(x)
  integer(8), intent(in) :: x
  minstd = mod(48271_8*x, 2147483647_8)
end function minstd

! Uniform (0,1) value of cell (i,j); identical integer recipe in the
! Python and Fortran case builders of the product.
real(8) function cell_uniform(sd, i, j)
  integer, intent(in) :: sd, i, j
  integer(8) :: x
  x = mod(int(sd,8)*1000003_8 + int(j,8)*7919_8 + int(i,8)*104729_8 + 12345_8, 2147483646_8) + 1_8
  x = minstd(x); x = minstd(x); x = minstd(x)
  cell_uniform = dble(x)/2147483647.d0
end function cell_uniform

! Cartesian Rayleigh-Taylor: heavy isothermal layer (T=1) over a light
! isothermal layer (T=2) with the density ratio 2 at the interface
! y_i(x) = 0.5*ylength + 0.01*cos(2*pi*x/xlength), gravity (0,-|phi|),
! and a seeded y-velocity noise of +-1e-4 per cell.
subroutine init_rt_cartesian(sd)
  integer, intent(in) :: sd
  real(8) :: yi, g, th, tl, rho, lam, vnoise, theta_down
  real(8), allocatable :: hh(:,:), bb(:,:)
  integer :: ii, jj
  g = abs(phi)
  phi = g
  th = 1.d0
  tl = 2.d0
  theta_down = atan2(-1.d0, 0.d0)
  allocate(hh(unum,vnum), bb(unum,vnum))
  do jj=iymin,iymax
    do ii=ixmin,ixmax
      yi = 0.5d0*ylength + 0.01d0*cos(2.d0*acos(-1.d0)*ctr(ii,jj)%x/xlength)
      if (ctr(ii,jj)%y >= yi) then
        rho = exp(-2.d0*g*(ctr(ii,jj)%y-yi)/th)
        lam = 1.d0/th
      else
        rho = 0.5d0*exp(2.d0*g*(yi-ctr(ii,jj)%y)/tl)
        lam = 1.d0/tl
      end if
      vnoise = 1.d-4*(2.d0*cell_uniform(sd, ii, jj)-1.d0)
      ctr(ii,jj)%prim = [rho, 0.d0, vnoise, lam]
      ctr(ii,jj)%w = get_conserved(ctr(ii,jj)%prim)
      ctr(ii,jj)%theta = theta_down
      call reduced_maxwell(hh, bb, uspace, vspace, ctr(ii,jj)%prim)
      ctr(ii,jj)%h = hh
      ctr(ii,jj)%b = bb
    end do
  end do
  do jj=iymin,iymax
    do ii=ixmin,ixmax+1
      vface(ii,jj)%theta = theta_down
    end do
  end do
  do jj=iymin,iymax+1
    do ii=ixmin,ixmax
      hface(ii,jj)%theta = theta_down
    end do
  end do
end subroutine init_rt_cartesian

end program ref_harness2d
