! Writes es23_16_values.txt: values in the reference's output format
! "(9(ES23.16,2X))" (1d/src/output.f90:95-97), including the edge cases of
! ESw.d output: signed zero, three-digit exponents, subnormals, an exact
! rounding tie at the 17th digit (2**-25), NaN and infinities.  Built and
! run by make_es_fixture.sh with the compiler of the reference build.
program es23_16_values
  use, intrinsic :: ieee_arithmetic
  implicit none
  real(8) :: v(40), z
  integer :: i, u
  z = 0.d0
  v(1) = 0.d0
  v(2) = -z
  v(3) = 1.d0
  v(4) = -1.d0
  v(5) = 0.1d0
  v(6) = 1.d0 / 3.d0
  v(7) = -2.d0 / 3.d0
  v(8) = 1.d-5
  v(9) = 123456.789d0
  v(10) = 1.d99
  v(11) = 9.999999999999999d99
  v(12) = 1.d100
  v(13) = -1.5d-100
  v(14) = 1.d-99
  v(15) = 1.d300
  v(16) = -1.d-300
  v(17) = tiny(1.d0)
  v(18) = huge(1.d0)
  v(19) = transfer(1_8, 1.d0)
  v(20) = 2.d0**(-25)
  v(21) = -2.d0**(-25)
  v(22) = 2.d0**(-29)
  v(23) = 0.5d0
  v(24) = 9.99999999999999999d-1
  v(25) = 99.99999999999999d0
  v(26) = 1.2345678901234567d-10
  v(27) = ieee_value(1.d0, ieee_quiet_nan)
  v(28) = ieee_value(1.d0, ieee_positive_inf)
  v(29) = ieee_value(1.d0, ieee_negative_inf)
  do i = 30, 40
    v(i) = sin(real(i, 8)) * 10.d0**(i - 35)
  end do
  open(newunit=u, file='es23_16_values.txt', status='replace', action='write')
  write(u, "(9(ES23.16,2X))") v
  write(u, "(9(ES23.16,2X))") v(1:0)
  write(u, "(9(ES23.16,2X))") v(1:9)
  close(u)
end program es23_16_values
