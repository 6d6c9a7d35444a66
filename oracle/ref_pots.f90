!-----------------------------------------------------------------------------
! ref_pots.f90 -- evaluates the reference's own Select_Pot (V_impur.f90:84-929)
! on a list of distances for a list of potential names.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile.ref with the
! unmodified V_impur.f90 where it lies; used to write golden vectors of the
! impurity potentials.  Nothing in the product links or calls it.
!
! stdin : n, then n distances (Angstrom), then one name per line
! out   : pots.bin (unformatted stream, float64): for each name, n values of
!         Select_Pot(name, r, r_cutoff = 0, umax = 0); then the He-TiO2
!         surface potentials of morse.f90: Morse_HeTiO2_1D(r) and
!         Morse_HeTiO2_3D(0.3 r, 1.7 - 0.2 r, r), n values each
!-----------------------------------------------------------------------------
program ref_pots
implicit none
double precision, external :: Select_Pot
real(8), external :: Morse_HeTiO2_1D, Morse_HeTiO2_3D
character (len=80) :: name
double precision, allocatable :: r(:)
double precision :: zero
integer :: n, i, ios
zero = 0.d0
read(5, *) n
allocate(r(n))
read(5, *) r
open(11, file='pots.bin', access='stream', form='unformatted', status='replace')
do
  read(5, '(A)', iostat=ios) name
  if (ios /= 0) exit
  if (len_trim(name) == 0) cycle
  do i = 1, n
    write(11) Select_Pot(name, r(i), zero, zero)
  end do
end do
do i = 1, n
  write(11) Morse_HeTiO2_1D(r(i))
end do
do i = 1, n
  write(11) Morse_HeTiO2_3D(0.3d0*r(i), 1.7d0 - 0.2d0*r(i), r(i))
end do
close(11)
end program ref_pots
