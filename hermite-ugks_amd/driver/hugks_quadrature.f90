! ------------------------------------------------------------
! Velocity quadratures built by the drivers at start-up.
!
! Gauss-Hermite: roots of the physicists' H_n by Newton iteration on the
! orthonormal three-term recurrence, started from the classical
! asymptotic guesses for the largest roots and extrapolated inward; the
! weight of a root is 2/p'_n^2 in that normalisation.  This is the
! published method the reference uses (2d/src/velocity.f90:63-137,
! 1d/src/velocity.f90:65-151).  The rule is returned, as there, as plain
! integration weights w*exp(x^2)*sqrt(Tref) at nodes x*sqrt(Tref).
!
! Newton-Cotes: closed 4-panel rule on [-vmax, vmax]*sqrt(Tref)
! (2d/src/velocity.f90:29-61).
! ------------------------------------------------------------
module hugks_quadrature
  implicit none
  private
  public :: gauss_hermite, newton_cotes_coeff

contains

  subroutine gauss_hermite(n, tref, pi, nodes, weights, roots, wgh)
    integer, intent(in) :: n
    real(8), intent(in) :: tref, pi
    real(8), intent(out) :: nodes(n), weights(n), roots(n), wgh(n)
    real(8), allocatable :: pos(:)
    real(8) :: z, zold, p0, p1, p2, dp, pstart
    integer :: i, j, it, m
    logical :: converged

    if (n < 2) error stop 'gauss_hermite: nv must be >= 2'
    m = (n + 1) / 2
    allocate(pos(m))
    pstart = pi**(-0.25d0)
    z = 0.d0
    dp = 1.d0
    do i = 1, m
      select case (i)
      case (1)
        z = sqrt(2.d0 * n + 1.d0) - 1.85575d0 * (2.d0 * n + 1.d0)**(-1.d0 / 6.d0)
      case (2)
        z = z - 1.14d0 * dble(n)**0.426d0 / z
      case (3)
        z = 1.86d0 * z - 0.86d0 * pos(1)
      case (4)
        z = 1.91d0 * z - 0.91d0 * pos(2)
      case default
        z = 2.d0 * z - pos(i - 2)
      end select
      converged = .false.
      do it = 1, 100
        p1 = pstart
        p2 = 0.d0
        do j = 1, n
          p0 = p2
          p2 = p1
          p1 = z * sqrt(2.d0 / j) * p2 - sqrt((j - 1.d0) / j) * p0
        end do
        dp = sqrt(2.d0 * n) * p2
        zold = z
        z = zold - p1 / dp
        if (abs(z - zold) <= 1.d-14 * (abs(z) + 1.d0)) then
          converged = .true.
          exit
        end if
      end do
      if (.not. converged) error stop 'gauss_hermite: root iteration did not converge'
      pos(i) = z
      roots(i) = -z
      roots(n + 1 - i) = z
      wgh(i) = 2.d0 / (dp * dp)
      wgh(n + 1 - i) = wgh(i)
    end do
    nodes = roots * sqrt(tref)
    do i = 1, n
      weights(i) = wgh(i) * exp(roots(i)**2) * sqrt(tref)
    end do
  end subroutine gauss_hermite

  real(8) function newton_cotes_coeff(idx, num)
    integer, intent(in) :: idx, num
    if (idx == 1 .or. idx == num) then
      newton_cotes_coeff = 14.d0 / 45.d0
    else if (mod(idx - 5, 4) == 0) then
      newton_cotes_coeff = 28.d0 / 45.d0
    else if (mod(idx - 3, 4) == 0) then
      newton_cotes_coeff = 24.d0 / 45.d0
    else
      newton_cotes_coeff = 64.d0 / 45.d0
    end if
  end function newton_cotes_coeff

end module hugks_quadrature
