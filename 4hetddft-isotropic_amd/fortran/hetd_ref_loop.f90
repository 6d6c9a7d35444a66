!-----------------------------------------------------------------------------
! hetd_ref_loop -- the reference's time loop written against module hetd_ref,
! routine by routine (DFT4He3d.f90:637-645 and 834-852): potenimp, poten,
! forceimp and aimp = F/(mimpur mp_u) after the state is read; steprk for the
! first three iterations and steppc after, each followed by potenimp, poten,
! forceimp and the new aimp.  Then energy() and derivnD of the final density
! and wave function along each axis.  It exercises the drop-in call surface
! as a program written against the reference would use it.
!
! Input: namelist &loop on stdin; the initial psi(nx,ny,nz), rimp(3) and
! vimp(3) as one raw stream file (statein).  Output, raw streams of real(8):
!   <prefix>traj.bin    per iteration errHe, errimp, rimp(3), vimp(3), aimp(3)
!   <prefix>energy.bin  etot4 ekin4 elj4 ecor4 esolid ealphas ekinx eHeX eimpu etot
!   <prefix>psi.bin     the final psi
!   <prefix>deriv.bin   d(den)/dx_iv, then d2(psi)/dx_iv**2, for iv = 1, 2, 3
!-----------------------------------------------------------------------------
program hetd_ref_loop
  use hetd_ref
  implicit none
  real(8), parameter :: mp_u = 0.02061484d0   ! modules.f90:315
  integer :: nx = 0, ny = 0, nz = 0, npd = 13, icon = 13, niter = 6, iter, iu, iv
  real(8) :: xmax = 0, ymax = 0, zmax = 0, r_cutoff = 2.0d0, umax = 0, mimpur = 131.29d0
  real(8) :: deltat = 0, tzmean = 1000.d0
  character(len=8) :: core4 = 'OT'
  character(len=80) :: selec = 'null'
  character(len=256) :: statein = 'state.bin', prefix = ''
  logical :: lsolid = .true.
  type(hetd_params) :: p
  complex(8), allocatable, target :: psi(:,:,:), dpsi(:,:,:)
  real(8), allocatable, target :: den(:,:,:), dden(:,:,:)
  real(8) :: rimp(3), vimp(3), aimp(3), F(3), errHe, errimp, h(3)
  namelist /loop/ nx, ny, nz, xmax, ymax, zmax, npd, icon, core4, lsolid, selec, r_cutoff, umax, &
                  mimpur, deltat, niter, tzmean, statein, prefix

  read(5, nml=loop)
  call hetd_default_params(p)
  p%nx = nx; p%ny = ny; p%nz = nz; p%npd = npd; p%icon = icon
  p%xmax = xmax; p%ymax = ymax; p%zmax = zmax
  p%core = merge(HETD_CORE_OTC, HETD_CORE_OT, core4 == 'OTC')
  p%lsolid = merge(1, 0, lsolid)
  p%mimpur = mimpur
  p%tzmean = tzmean
  call hetd_setup(p, selec, r_cutoff, umax)
  call fftini(nx, ny, nz, 1)

  allocate(psi(nx,ny,nz), dpsi(nx,ny,nz), den(nx,ny,nz), dden(nx,ny,nz))
  open(newunit=iu, file=statein, access='stream', form='unformatted', status='old')
  read(iu) psi, rimp, vimp
  close(iu)
  call hetd_check(hetd_set_state(hctx, psi, rimp, vimp), 'hetd_set_state')

  ! DFT4He3d.f90:641-645
  call potenimp(rimp)
  call poten()
  call forceimp(rimp, F)
  aimp = F/(mimpur*mp_u)
  call hetd_check(hetd_set_aimp(hctx, aimp), 'hetd_set_aimp')

  open(newunit=iu, file=trim(prefix)//'traj.bin', access='stream', form='unformatted', status='replace')
  do iter = 1, niter
    ! DFT4He3d.f90:840-852
    if (iter <= 3) then
      call steprk(deltat)
      errHe = 0.d0
      errimp = 0.d0
    else
      call steppc(deltat, errHe, errimp)
    end if
    call hetd_check(hetd_get_imp(hctx, rimp, vimp, aimp), 'hetd_get_imp')
    call potenimp(rimp)
    call poten()
    call forceimp(rimp, F)
    aimp = F/(mimpur*mp_u)
    call hetd_check(hetd_set_aimp(hctx, aimp), 'hetd_set_aimp')
    write(iu) errHe, errimp, rimp, vimp, aimp
  end do
  close(iu)

  call energy()
  open(newunit=iu, file=trim(prefix)//'energy.bin', access='stream', form='unformatted', status='replace')
  write(iu) etot4, ekin4, elj4, ecor4, esolid, ealphas, ekinx, eHeX, eimpu, etot
  close(iu)

  call hetd_check(hetd_get_field(hctx, HETD_F_PSI, c_loc(psi)), 'hetd_get_field')
  open(newunit=iu, file=trim(prefix)//'psi.bin', access='stream', form='unformatted', status='replace')
  write(iu) psi
  close(iu)

  ! derden (derden.f90:26-28) and the kinetic Laplacian terms, through derivnD
  den = real(psi)**2 + aimag(psi)**2
  h = 2.0d0*abs([xmax, ymax, zmax])/[nx, ny, nz]
  open(newunit=iu, file=trim(prefix)//'deriv.bin', access='stream', form='unformatted', status='replace')
  do iv = 1, 3
    call derivnD(1, [nx, ny, nz], h(iv), iv, den, dden, icon)
    write(iu) dden
  end do
  do iv = 1, 3
    call derivnD(2, [nx, ny, nz], h(iv), iv, psi, dpsi, icon)
    write(iu) dpsi
  end do
  close(iu)
  call hetd_check(hetd_finalize(hctx), 'hetd_finalize')
end program hetd_ref_loop
