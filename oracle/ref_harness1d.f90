! ------------------------------------------------------------------
! ref_harness1d: test-infrastructure driver around the reference 1D
! solver modules (compiled unmodified from the reference sources by
! oracle/Makefile).  FP64 oracle and CPU baseline only.
!
! Replaces only the reference main program (1d/src/main.f90:14-139):
! exit test at the top of the loop (1d/src/main.f90:84), phases
! timestep/interpolation/evolution/update (86-89), iter += 1 then
! simtime += dt (106-107), plus a step cap.  The Fourier steady-state
! exit (109-126) is reproduced with the reference's own predicate.
!
! usage: ref_harness1d <namelist> <dump|-> <max_steps>
!
! dump layout (stream, native endian):
!   int32 nx, nv, steps, iter
!   real64 simtime, dt, res(3), steady_res(3)
!   real64 dts(steps)
!   real64 u(nv), weight(nv)
!   real64 w(3, nx), h(nv, nx), b(nv, nx)
! ------------------------------------------------------------------
program ref_harness1d
use CONFIG
use ALLDATA
use GRIDMAKER
use INITIALIZER
use SOLVER
implicit none

character(len=256) :: cfgfile, dumpfile, arg
integer :: max_steps, steps, i, u, steady_count
integer(8) :: c0, c1, crate, tphase(4), tloop0, tloop1
real(8), allocatable :: dts(:)
real(8) :: mass0, mass1, rel_mass

call get_command_argument(1, cfgfile)
call get_command_argument(2, dumpfile)
call get_command_argument(3, arg); read(arg,*) max_steps

call read_config(trim(cfgfile))
isRestart = .false.
iniType = 1
gridfile = 'none'
call gridInit()
call initialize()

allocate(dts(max(max_steps,1)))
mass0 = 0.d0
do i=ixmin,ixmax
  mass0 = mass0 + ctr(i)%w(1)*ctr(i)%length
end do

call system_clock(count_rate=crate)
tphase = 0
steps = 0
steady_count = 0
call system_clock(tloop0)
do while (steps < max_steps)
  if (simtime >= maxtime) exit
  call system_clock(c0); call timestep();      call system_clock(c1); tphase(1)=tphase(1)+c1-c0
  call system_clock(c0); call interpolation(); call system_clock(c1); tphase(2)=tphase(2)+c1-c0
  call system_clock(c0); call evolution();     call system_clock(c1); tphase(3)=tphase(3)+c1-c0
  call system_clock(c0); call update();        call system_clock(c1); tphase(4)=tphase(4)+c1-c0
  steps = steps + 1
  dts(steps) = dt
  iter = iter + 1
  simtime = simtime + dt
  if (case_id == 4) then
    mass1 = 0.d0
    do i=ixmin,ixmax
      mass1 = mass1 + ctr(i)%w(1)*ctr(i)%length
    end do
    rel_mass = abs(mass1-mass0)/(abs(mass0)+SMV)
    if (fourier_is_steady(steady_res, rel_mass, steady_tol, mass_tol)) then
      steady_count = steady_count + 1
    else
      steady_count = 0
    end if
    if (steady_count >= steady_steps) exit
  end if
end do
call system_clock(tloop1)

mass1 = 0.d0
do i=ixmin,ixmax
  mass1 = mass1 + ctr(i)%w(1)*ctr(i)%length
end do
write(*,'(A,I8,A,I8,A,ES24.16,A,ES24.16)') 'HARNESS steps=', steps, ' iter=', iter, &
     ' simtime=', simtime, ' rel_mass_change=', (mass1-mass0)/mass0
write(*,'(A,ES14.6,A,4ES14.6)') 'HARNESS_TIMING loop_s=', dble(tloop1-tloop0)/dble(crate), &
     ' phases_s=', dble(tphase)/dble(crate)

if (trim(dumpfile) /= '-') then
  open(newunit=u, file=trim(dumpfile), access='stream', form='unformatted', status='replace')
  write(u) int(nx,4), int(nv,4), int(steps,4), int(iter,4)
  write(u) simtime, dt, res, steady_res
  write(u) dts(1:steps)
  write(u) uspace, weight
  do i=ixmin,ixmax
    write(u) ctr(i)%w
  end do
  do i=ixmin,ixmax
    write(u) ctr(i)%h
  end do
  do i=ixmin,ixmax
    write(u) ctr(i)%b
  end do
  close(u)
end if

end program ref_harness1d
