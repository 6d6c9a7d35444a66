!-----------------------------------------------------------------------------
! ref_harness.f90 -- test harness around the reference's own Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  This program is compiled (by oracle/Makefile.ref)
! together with the unmodified reference sources where they lie, and is used
! to write golden vectors and to time the reference CPU path.  Nothing in the
! product links or calls it.
!
! It replays the reference driver's set-up and time loop for the real-time
! propagation step, calling the reference routines themselves:
!   set-up order      DFT4He3d.f90:171-832  (grids :274-361, potenimpini :373,
!                     readenc mode 2 :379, core selection :476-509, fftini :518,
!                     fforma :526, initcg :579-590, kalfs :592-596,
!                     first potenimp/poten/forceimp/energy :637-645,
!                     timec :689-699, impurity histories :796-802)
!   time loop         DFT4He3d.f90:834-852 (RK for the first 3 steps, then PC,
!                     then potenimp/poten/forceimp at the corrected state)
!   diagnostics       DFT4He3d.f90:866-872 (norm, energy every pener steps)
! The input state is a mode-2 file (readenc.f90:104-113) so that this harness
! and the GPU path read the same bytes.
!
! Output (unformatted stream, little-endian float64) into outdir:
!   init.bin   : fields after the first poten/forceimp/energy
!   steps.bin  : per step  [iter, errHe, errimp, rimp(3), vimp(3), aimp(3)]
!   energy.bin : every pener steps [iter, norm, etot4, ekin4, elj4, ecor4,
!                esolid, ealphas, ekinx, eHeX, etot]
!   final.bin  : psi after niter steps (complex128, x fastest)
!   timing.txt : wall seconds of the time loop and the number of steps
!-----------------------------------------------------------------------------
program ref_harness
use seleccio_de_potencial
use alphasterm
use Para_DerivnD
use deriva
use energies
use rho
use field
use fftmodule
use grid
use gridk
use classicimp
use rkpc
use impur, only: r_cutoff, selec, umax
use lenard4
use he4
use util1
use work1
implicit none

integer (kind=4) :: ix, iy, iz, ipx, ipy, ipz, iter, niter, pener, nthread
integer (kind=4) :: ndmax, mode, n4, iteraux, it0, nrk
integer (kind=8) :: t_1, t_2, irate
real    (kind=8) :: deltat, deltatps, mimpur, rimpur, r_clust, p, px2, py2, pz2
real    (kind=8) :: errHe, errimp, auxn4, aux
real    (kind=8) :: Lambdah, txmean, txsurf, tymean, tysurf, tzmean, tzsurf
character (len=60)  :: filedenin, fileimpin
character (len=200) :: outdir
logical :: ldump_init

namelist /harness/ nx, ny, nz, xmax, ymax, zmax, npd, icon, core4, lsolid,  &
                   selec, r_cutoff, umax, mimpur, deltatps, niter, pener,   &
                   filedenin, outdir, Lambdah, txmean, txsurf, tymean,      &
                   tysurf, tzmean, tzsurf, nthread, ldump_init, n4, mode, nrk

! defaults mirror the reference driver's local declarations
niter = 10; pener = 5; nthread = 1; mimpur = 131.29d0; deltatps = 5.d-4
Lambdah = 2.d0; txmean = 1000.d0; txsurf = 2.d0; tymean = 1000.d0
tysurf = 2.d0; tzmean = 1000.d0; tzsurf = 2.d0; ldump_init = .true.
filedenin = 'den.in'; fileimpin = 'none'; outdir = '.'; n4 = 1000
selec = 'null'; r_cutoff = 2.d0; umax = 0.d0
rimpur = 5.5d0; mode = 2; ndmax = 2; nrk = 3
! FFTW_ESTIMATE: the reference keeps a CHARACTER plan name where the FFTW API
! expects INTEGER flags (fft.f90:50-64); the first four bytes are what the
! library reads, so they are set to the ESTIMATE flag value (64).
fftwplan = char(64)//char(0)//char(0)//char(0)

read(5, nml=harness)

pi = 4.0d0*datan(1.0d0); twopi = 2.0d0*pi; fourpi = 4.0d0*pi; piq = pi*pi
selec_gs = selec; r_cutoff_gs = r_cutoff; umax_gs = umax
mAg_u = mimpur*mp_u
nn(1) = nx; nn(2) = ny; nn(3) = nz
mmx(1) = nx; mmx(2) = ny; mmx(3) = nx; mmx(4) = ny
deltat = deltatps/7.63822291d0

call Init_deriv_p(npd, ndmax, nthread)

hx = 2.0d0*abs(xmax)/(nx); hy = 2.0d0*abs(ymax)/(ny); hz = 2.0d0*abs(zmax)/(nz)
dxyz = hx*hy*hz; nxyz = nx*ny*nz; dVomAg = dxyz/mAg_u
hpx = 1.0d0/(nx*hx); hpy = 1.0d0/(ny*hy); hpz = 1.0d0/(nz*hz)
pmaxx = 1.0d0/(2.0d0*hx); pmaxy = 1.0d0/(2.0d0*hy); pmaxz = 1.0d0/(2.0d0*hz)

call dimen()

rimp = 0.d0; vimp = 0.d0
do ix = 1, nx; x(ix) = -xmax + hx*(ix-1); end do
do iy = 1, ny; y(iy) = -ymax + hy*(iy-1); end do
do iz = 1, nz; z(iz) = -zmax + hz*(iz-1); end do
do ipx = 1, nx/2+1;    px(ipx) = hpx*(ipx-1); end do
do ipx = nx/2+2, nx;   px(ipx) = -pmaxx + hpx*(ipx-(nx/2+1)); end do
do ipy = 1, ny/2+1;    py(ipy) = hpy*(ipy-1); end do
do ipy = ny/2+2, ny;   py(ipy) = -pmaxy + hpy*(ipy-(ny/2+1)); end do
do ipz = 1, nz/2+1;    pz(ipz) = hpz*(ipz-1); end do
do ipz = nz/2+2, nz;   pz(ipz) = -pmaxz + hpz*(ipz-(nz/2+1)); end do
do ipz = 1, nz
  pz2 = pz(ipz)**2
  do ipy = 1, ny
    py2 = py(ipy)**2
    do ipx = 1, nx/2+1
      px2 = px(ipx)**2
      pmod(ipx,ipy,ipz) = sqrt(px2+py2+pz2)
    end do
  end do
end do

call potenimpini()
if (mode .eq. 2) then
  call readenc(n4, densat4, filedenin, fileimpin, mode, rimpur, r_clust)
else
  ! timing runs only: Woods-Saxon droplet built in place (no input file),
  ! rho0=0.0218, R=2.22 N^(1/3), a=0.7, a small deterministic phase
  do iz = 1, nz
    do iy = 1, ny
      do ix = 1, nx
        aux = sqrt(x(ix)**2 + y(iy)**2 + z(iz)**2)
        aux = 0.0218d0/(1.d0 + exp(min((aux - 2.22d0*dble(n4)**(1.d0/3.d0))/0.7d0, 700.d0)))
        psi(ix,iy,iz) = sqrt(aux)*cmplx(cos(1.d-3*sin(0.3d0*x(ix)+0.2d0*y(iy))),   &
                                        sin(1.d-3*sin(0.3d0*x(ix)+0.2d0*y(iy))), kind=8)
      end do
    end do
  end do
  rimp = 0.d0; vimp = 0.d0; vimp(3) = 2.d0*7.63822291d0
end if

if (lsolid) core4 = 'OT '
select case (core4)
  case ('OT ')
    h4 = h4ot
  case ('OTC')
    h4 = h4ot
    allocate(denalf(nx,ny,nz)); allocate(falfs(nx,ny,nz))
    allocate(kalfs(nx/2+1,ny,nz)); allocate(intxalf(nx,ny,nz))
    allocate(intyalf(nx,ny,nz)); allocate(intzalf(nx,ny,nz))
    allocate(ualphas(nx,ny,nz))
  case default
    stop 'harness: unsupported core4'
end select

call fftini(nx, ny, nz, nthread)
call fforma(core4, b4, eps4, sigma4, h4, nx, ny, nz, pmod, fvlj4)
if (lsolid) then
  aux = h4*1.065d0
  call initcg(aux, wcgk)
else
  call initcg(h4, wcgk)
end if
if (core4 .eq. 'OTC') then
  forall (ix=1:nx/2+1, iy=1:ny, iz=1:nz)
    kalfs(ix,iy,iz) = exp(-(pi*l*pmod(ix,iy,iz))**2)
  end forall
end if

den = conjg(psi)*psi
call potenimp(rimp)
call poten()
call forceimp(rimp, aimp)
aimp(:) = aimp(:)/mAg_u
uext = 0.d0
call energy()

if (ldump_init) then
  open(21, file=trim(outdir)//'/init.bin', access='stream', form='unformatted', status='replace')
  write(21) pot4, dencg, delj4, sto1, dxden, dyden, dzden, uimp, fvlj4, wcgk, fden
  if (lsolid) write(21) penalty
  if (core4 .eq. 'OTC') write(21) ualphas, denalf
  write(21) etot4, ekin4, elj4, ecor4, esolid, ealphas, ekinx, eHeX, etot, aimp
  close(21)
end if

do iz = 1, nz
  do iy = 1, ny
    do ix = 1, nx
      timec(ix,iy,iz) = cmplx(Lambdah*(1.d0+tanh((abs(x(ix))-txmean)/txsurf)      &
                              +1.d0+tanh((abs(y(iy))-tymean)/tysurf)              &
                              +1.d0+tanh((abs(z(iz))-tzmean)/tzsurf)), 1.d0, kind=8)
    end do
  end do
end do
uext = 0.d0
qr = 0.d0; qv = 0.d0
rimpold = 0.d0; vimpold = 0.d0; aimpold = 0.d0

open(22, file=trim(outdir)//'/steps.bin',  access='stream', form='unformatted', status='replace')
open(23, file=trim(outdir)//'/energy.bin', access='stream', form='unformatted', status='replace')

it0 = 1
call system_clock(count=t_1, count_rate=irate)
do iter = it0, niter
  iteraux = iter - it0 + 1
  errHe = 0.d0; errimp = 0.d0
  if (iteraux .eq. nrk + 1) call system_clock(count=t_1)
  if (iteraux .le. nrk) then
    call steprk(deltat)
  else
    call steppc(deltat, errHe, errimp)
  end if
  call potenimp(rimp)
  call poten()
  call forceimp(rimp, aimp)
  aimp(:) = aimp(:)/mAg_u
  write(22) dble(iter), errHe, errimp, rimp, vimp, aimp
  if (mod(iter, pener) .eq. 0) then
    auxn4 = sum(den)*dxyz
    call energy()
    write(23) dble(iter), auxn4, etot4, ekin4, elj4, ecor4, esolid, ealphas, ekinx, eHeX, etot
  end if
end do
call system_clock(count=t_2)
close(22); close(23)

open(24, file=trim(outdir)//'/final.bin', access='stream', form='unformatted', status='replace')
write(24) psi
close(24)
open(25, file=trim(outdir)//'/timing.txt', status='replace')
write(25, '(E20.10, I10)') dble(t_2 - t_1)/dble(irate), niter - min(nrk, niter)
close(25)
end program ref_harness
