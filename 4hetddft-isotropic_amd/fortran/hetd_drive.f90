!-----------------------------------------------------------------------------
! hetd_drive -- driver of the real-time propagation on one MI355X.
!
! Reads the reference's namelist &input from stdin (DFT4He3d.f90:149-168;
! every key is accepted, the ones that do not concern this path are read and
! ignored) and a state file of mode 0, 2 or 3 (readenc.f90:80-129), then runs the
! reference loop (DFT4He3d.f90:834-1004): Runge-Kutta-Gill for the first
! three iterations (or always with lrk), Hamming predictor-corrector after,
! the mean field and impurity force at the new state, the impurity R/v/a
! files every step, energies every pener steps and a checkpoint in the
! printoutc layout (printoutc.f90:52-76) every pcurr steps and at the end.
! The iterations between two outputs run on the device without host
! synchronisation (hetd_evolve); with lrk the per-routine entry points are
! called one by one.
!-----------------------------------------------------------------------------
program hetd_drive
  use hetd_capi
  use hetd_ref
  implicit none

  ! namelist variables (names and defaults of DFT4He3d.f90:52-145, modules.f90)
  character(len=40) :: title = 'Helium4 - 3dim.  '
  character(len=15) :: fftwplan = 'FFTW_PATIENT'
  integer :: nthread = 1, nsfiles = 1, n4 = 300, mode = 0
  character(len=60) :: fileout = 'DFT4He3d.res', filemonopol = 'monopol.out'
  character(len=60) :: filerimp = 'rimp.out', filevimp = 'vimp.out', fileaimp = 'aimp.out'
  character(len=60) :: filedenin = 'he4.input.density', filedenout = 'he4.output.density'
  character(len=60) :: fileimpin = 'X.input.wf', fileimpout = 'X.output.wf'
  integer :: nx = 64, ny = 64, nz = 64
  real(8) :: xmax = 64.d0, ymax = 64.d0, zmax = 64.d0, xc = 0, yc = 0, zc = 0, afermi = 0
  real(8) :: eps4 = 10.22d0, sigma4 = 2.556d0, l = 1.0d0
  character(len=3) :: core4 = 'OT '
  real(8) :: cp4 = -2.41186d4, cpp4 = 1.85850d6, den4c = 0.062d0, alphas = 54.31d0, h2o2m4 = 6.05969638298d0
  real(8) :: denmin = 1.d-60, psimin = 1.d-30
  integer :: npd = 13, ndmax = 2, icon = 13, niter = 0, pchem = 50, irespar = 0
  logical :: printpot = .false., lsolid = .true.
  integer :: pdenpar = 50, pener = 50, ppot = 1
  real(8) :: ximp = 0, yimp = 0, zimp = 0, rimpur = 5.5d0, mimpur = 0
  real(8) :: deltat = 0.d0, deltatps = 0.d0, xlamda = 0, xlamdax = 0, xlamday = 0, xlamdaz = 0
  logical :: lrkpc = .true., lrk = .false., lselection = .true.
  real(8) :: xlamdax_x = 0, xlamdax_y = 0, xlamdax_z = 0, time0 = -1, Zsurf = -25.d0
  integer :: iter0 = 0, pcurr = 0, icurr = 0
  logical :: lsurf = .false., lsurf3D = .false.
  real(8) :: Lambdah = 2.d0, txmean = 1000.d0, txsurf = 2.d0, tymean = 1000.d0, tysurf = 2.d0
  real(8) :: tzmean = 1000.d0, tzsurf = 2.d0
  real(8) :: vximp = 0, vyimp = 0, vzimp = 0
  character(len=80) :: selec = '', selec_gs = ''
  real(8) :: r_cutoff = 2.0d0, umax = 400.d0, r_cutoff_gs = -1.d0, umax_gs = -1.d0
  ! extension (not a reference key): keep the integrator history beside every
  ! checkpoint (<file>.hist) and restart from <filedenin>.hist when present
  logical :: lhistory = .false.
  ! extension (not a reference key): run the mixed-precision path (FP32 mean
  ! field and FFTs, FP64 wave function and integrator state)
  logical :: lmixed = .false.

  namelist /input/title,fftwplan,nthread,nsfiles,                         &
                  fileout,filemonopol, filerimp,filevimp,fileaimp,        &
                  filedenin,filedenout,                                   &
                  fileimpin,fileimpout,                                   &
                  n4,mode,                                                &
                  nx,ny,nz,xmax,ymax,zmax,                                &
                  xc,yc,zc,afermi,                                        &
                  eps4,sigma4,core4,l,                                    &
                  cp4,cpp4,den4c,alphas,h2o2m4,                           &
                  denmin,psimin,npd,ndmax,icon,                           &
                  niter,printpot,pchem,irespar,                           &
                  pdenpar,pener,ppot,lsolid,                              &
                  ximp,yimp,zimp,rimpur,mimpur,                           &
                  deltat,lrkpc,lrk,xlamda,lselection,                     &
                  xlamdax,xlamday,xlamdaz,deltatps,                       &
                  xlamdax_x,xlamdax_y,xlamdax_z,iter0,time0,Zsurf,        &
                  pcurr,icurr,lsurf,lsurf3D,Lambdah,                      &
                  txmean,txsurf,tymean,tysurf,tzmean,tzsurf,              &
                  vximp,vyimp,vzimp,selec,r_cutoff,umax,                  &
                  selec_gs,r_cutoff_gs,umax_gs,lhistory,lmixed

  real(8), parameter :: ktops = 7.63822291d0, pstok = 1.d0/7.63822291d0, mp_u = 0.02061484d0
  type(hetd_params) :: p
  complex(8), allocatable, target :: psi(:,:,:)
  real(8), allocatable :: traj(:,:)
  real(8) :: rimp(3), vimp(3), aimp(3), mAg_u, auxn4, eold, temps
  real(8) :: hx, hy, hz, t_start, t_end
  real(8) :: cm(3), r2(3), ang(3)
  integer :: iter, iteraux, nchunk, s, next_out, ncurr
  integer(c_int) :: rc
  integer(8) :: c1, c2, crate
  logical :: use_evolve

  read(5, nml=input, end=999)
  if (selec_gs == '') selec_gs = selec
  if (r_cutoff_gs < 0.d0) r_cutoff_gs = r_cutoff
  if (umax_gs < 0.d0) umax_gs = umax
  mAg_u = mimpur*mp_u
  if (tstgrid(nx) /= 0) error stop 'SEVERE ERROR: NX is not correct'
  if (tstgrid(ny) /= 0) error stop 'SEVERE ERROR: NY is not correct'
  if (tstgrid(nz) /= 0) error stop 'SEVERE ERROR: NZ is not correct'
  ! deltat consistency (DFT4He3d.f90:233-249)
  if (deltat == 0.d0) then
    if (deltatps == 0.d0) error stop 'You must specify either Deltat (in kelvin) or Deltatps (in picosecond)'
    deltat = deltatps/ktops
  else
    if (deltatps == 0.d0) then
      deltatps = deltat*ktops
    else if (deltatps /= deltat*ktops) then
      error stop 'Inconsistent deltat - deltatps'
    end if
  end if
  ! readenc.f90:136-140 (modes 0, 2 and 3 are read)
  if (mode /= 0 .and. mode /= 2 .and. mode /= 3) error stop 'ERROR READEN: This mode is still not programmed'
  if (lsolid) core4 = 'OT '
  if (core4 /= 'OT ' .and. core4 /= 'OTC') error stop ' ... The program stops due to a severe error.'

  hx = 2.0d0*abs(xmax)/nx; hy = 2.0d0*abs(ymax)/ny; hz = 2.0d0*abs(zmax)/nz

  call hetd_default_params(p)
  p%nx = nx; p%ny = ny; p%nz = nz; p%npd = npd; p%icon = icon
  p%core = merge(HETD_CORE_OTC, HETD_CORE_OT, core4 == 'OTC')
  p%lsolid = merge(1, 0, lsolid); p%precision = merge(HETD_MIXED, HETD_FP64, lmixed)
  p%xmax = xmax; p%ymax = ymax; p%zmax = zmax
  p%eps4 = eps4; p%sigma4 = sigma4; p%cp4 = cp4; p%cpp4 = cpp4; p%h2o2m4 = h2o2m4
  p%alphas = alphas; p%l_alpha = l
  p%mimpur = mimpur
  p%Lambdah = Lambdah; p%txmean = txmean; p%txsurf = txsurf; p%tymean = tymean
  p%tysurf = tysurf; p%tzmean = tzmean; p%tzsurf = tzsurf
  call hetd_setup(p, selec_gs, r_cutoff_gs, umax_gs)
  call fftini(nx, ny, nz, nthread)

  ! the impurity of the namelist (DFT4He3d.f90:300-302); modes 2 and 3 read it from the file
  rimp = (/ximp, yimp, zimp/)
  vimp = (/vximp, vyimp, vzimp/)*ktops
  allocate(psi(nx,ny,nz))
  if (mode == 0) then
    call read_mode0(filedenin)
  else
    call read_mode23(filedenin)
  end if
  rc = hetd_set_state(hctx, psi, rimp, vimp)
  call hetd_check(rc, 'hetd_set_state')

  ! external potential of the TiO2 surface (DFT4He3d.f90:758-784): the
  ! reference adds it to the energy only
  if (lsurf) then
    write(6, '(" Compute surface potential")')
    call hetd_check(hetd_set_surface(hctx, merge(1, 0, lsurf3D), Zsurf), 'set_surface')
  end if
  write(6, '(" Time step is ",1p,E16.8," kelvins or ",E16.8," picoseconds.")') deltat, deltatps
  ! first potential and energy (DFT4He3d.f90:641-645)
  call potenimp(rimp)
  call poten()
  call forceimp(rimp, aimp)
  aimp = aimp/mAg_u
  call hetd_check(hetd_set_aimp(hctx, aimp), 'set_aimp')
  call energy()
  call hetd_check(hetd_norm(hctx, auxn4), 'norm')
  write(6, '("Number of He4 atoms",1P,E15.6)') auxn4
  write(6, 6050) etot4, etot4/n4, ekin4, elj4, ealphas, esolid, ecor4
  write(6, 6060) eimpu, ekinx, eHeX, 0.d0, etot
  eold = etot4
  if (lhistory) call read_history(trim(filedenin)//'.hist')

  open(81, file=filerimp); open(82, file=filevimp); open(83, file=fileaimp)
  write(81, '("# Tiempo(ps), x(AA), y(AA), z(AA)")')
  write(81, '(1x,1p,E15.6,3E18.10)') time0, rimp
  write(82, '("# Tiempo(ps), Vx(AA/ps), Vy(AA/ps), Vz(AA/ps)")')
  write(82, '(1x,1p,E15.6,3E18.10)') time0, vimp*pstok
  write(83, '("# Tiempo(ps), Ax(AA/ps**2), Ay(AA/ps**2), Az(AA/ps**2)")')
  write(83, '(1x,1p,E15.6,3E18.10)') time0, aimp*pstok*pstok

  use_evolve = .not. lrk
  if (pcurr <= 0) pcurr = 1000000000
  if (pener <= 0) pener = 1000000000
  allocate(traj(HETD_TRAJ_STRIDE, max(1, min(niter - iter0, 4096))))
  call system_clock(count=c1, count_rate=crate)
  iter = iter0
  do while (iter < niter)
    ! run up to the next energy / checkpoint iteration on the device
    next_out = min(niter, ((iter/pener) + 1)*pener, iter0 + (((iter - iter0)/pcurr) + 1)*pcurr)
    nchunk = min(next_out - iter, size(traj, 2))
    if (use_evolve) then
      rc = hetd_evolve(hctx, deltat, int(nchunk, c_int32_t), traj)
      call hetd_check(rc, 'hetd_evolve')
    else
      do s = 1, nchunk
        call steprk(deltat)
        call hetd_check(hetd_get_imp(hctx, rimp, vimp, aimp), 'get_imp')
        call potenimp(rimp)
        call poten()
        call forceimp(rimp, aimp)
        aimp = aimp/mAg_u
        call hetd_check(hetd_set_aimp(hctx, aimp), 'set_aimp')
        traj(1:2, s) = 0.d0
        traj(3:5, s) = rimp; traj(6:8, s) = vimp; traj(9:11, s) = aimp
      end do
    end if
    do s = 1, nchunk
      iter = iter + 1
      iteraux = iter - iter0
      if (iteraux > 3 .and. use_evolve) &
        write(6, '(" Error( He, imp) (From Steppc)...",1p,2E15.6)') traj(1, s), traj(2, s)
      temps = time0 + iteraux*deltatps
      write(81, '(1x,1p,E15.6,3E18.10)') temps, traj(3:5, s)
      write(82, '(1x,1p,E15.6,3E18.10)') temps, traj(6:8, s)*pstok
      write(83, '(1x,1p,E15.6,3E18.10)') temps, traj(9:11, s)*pstok*pstok
    end do
    flush(81); flush(82); flush(83)
    call hetd_check(hetd_get_imp(hctx, rimp, vimp, aimp), 'get_imp')
    if (mod(iter, pener) == 0) then
      write(6, '(" Iteration....:",I10)') iter
      call hetd_check(hetd_norm(hctx, auxn4), 'norm')
      write(6, '(" Number of particles....:",1p,E18.10)') auxn4
      call energy()
      write(6, 7010) etot4, (etot4 - eold), etot4/n4, ekin4, elj4, ealphas, esolid, ecor4
      write(6, 7015) eimpu, ekinx, eHeX, 0.d0, etot
      eold = etot4
      call moments()
      write(6, '("<Lx,Ly,Lz>.......:",1p,3E20.11)') ang
      write(6, '(/1x,T5,"Center of Mass of the Helium ...(",0P,F9.4,",",F9.4,",",F9.4,") A")') cm
      write(6, '(" Impurity position:",1p,3E15.6)') rimp
      write(6, '(" Impurity velocity:",1p,3E15.6)') vimp
    end if
    if (mod(iter - iter0, pcurr) == 0) then
      ncurr = (iter - iter0)/pcurr + icurr
      call write_mode2(density_name(ncurr), iter, time0 + (iter - iter0)*deltatps)
    end if
  end do
  call system_clock(count=c2)
  t_end = dble(c2 - c1)/dble(crate)
  t_start = 0
  print *, "elapsed partie iterative : ", real(t_end - t_start)
  call write_mode2(filedenout, iter, time0 + (iter - iter0)*deltatps)
  call hetd_check(hetd_finalize(hctx), 'finalize')
  stop
999 error stop 'DFT3He3d. Error in input master file. Too short'

6050 format(//,T5,'FIRST ENERGY BALANCE: ',                    //     &
              ,T5,'TOTAL   energy (He) ..........: ',F18.6,' K',/,    &
               T5,'Energy per particle (He) .....: ',F18.6,' K',/,    &
               T5,'Kinetic energy (He) ..........: ',F18.6,' K',/,    &
               T5,'Lennard-Jones energy (He) ....: ',F18.6,' K',/,    &
               T5,'Alpha_s term  energy (He) ....: ',F18.6,' K',/,    &
               T5,'Solid energy (He)  ...........: ',F18.6,' K',/,    &
               T5,'Correlation energy   (He) ....: ',F18.6,' K')
6060 format(1x,T5,'Impurity energy (X) ..........: ',F18.6,' K',/,    &
               T5,'Kinetic energy (X) ...........: ',F18.6,' K',/,    &
               T5,'Interaction energy (X-He) ....: ',F18.6,' K',/,    &
               T5,'Spin-Orbit energy (X) ........: ',F18.6,' K',/,    &
               T5,'TOTAL ENERGY (He+X) ..........: ',F18.6,' K',/)
7010 format(//,T5,'ITERATIVE PROCEDURE ',                                    //  &
              ,T5,'Total Energy (He).......... ',0P,F18.6,' K +- ',1P,e12.4,' K',&
             /,T5,'Energy per particle (He)... ',0P,F18.6,' K',/,                &
             /,T5,'Kinetic Energy (He)........ ',0P,F18.6,' K',                  &
             /,T5,'Lennard-Jones Energy (He).. ',0P,F18.6,' K',                  &
             /,T5,'Alpha_s term  energy (He).. ',0P,F18.6,' K',                  &
             /,T5,'Solid energy (He) ......... ',0P,F18.6,' K',                  &
             /,T5,'Correlation Energy  (He)... ',0P,F18.6,' K')
7015 format(   T5,'Impurity energy (X->He) ... ',0P,F18.6,' K',/,&
               T5,'Kinetic energy (X) ........ ',0P,F18.6,' K',/,&
               T5,'Interaction energy (X-He) . ',0P,F18.6,' K',/,    &
               T5,'Spin-Orbit energy (X) ..... ',0P,F18.6,' K',/,    &
               T5,'TOTAL energy (He+X) ....... ',0P,F18.6,' K',/,' ')

contains

  integer function tstgrid(n)
    ! grid sizes accepted by the reference (tstgrid.f90:8-12)
    integer, intent(in) :: n
    integer, parameter :: nvalid(53) = (/1,2,3,4,5,6,8,9,10,12,15,16,18,20,24,25,27,30,32,36, &
        40,45,48,50,54,60,64,72,75,80,81,90,96,100,108,120,125,128,135,144,150,160,162,180,  &
        192,200,216,225,240,243,250,256,320/)
    tstgrid = 1
    if (any(nvalid == n)) tstgrid = 0
  end function tstgrid

  function density_name(k) result(name)
    integer, intent(in) :: k
    character(len=60) :: name
    character(len=4) :: chariter
    write(chariter, '(I4.4)') k
    name = 'density.'//chariter//'.dat'
  end function density_name

  subroutine titols(iu)
    ! skip the leading '#' comment lines of a state file (titols.f90:12-42)
    integer, intent(in) :: iu
    character(len=1) :: c
    integer :: ios
    do
      read(iu, '(A1)', iostat=ios) c
      if (ios /= 0) error stop 'STOP due to severe errors in routine TITOLS'
      if (c /= '#') exit
    end do
    backspace(iu)
  end subroutine titols

  subroutine check_grid(nxp, nyp, nzp)
    integer, intent(in) :: nxp, nyp, nzp
    if (nxp /= nx .or. nyp /= ny .or. nzp /= nz) error stop 'hetd_drive: state file grid differs from the namelist'
  end subroutine check_grid

  subroutine read_mode23(fname)
    ! readenc modes 2 and 3 (readenc.f90:104-129): grid line, rimp, vimp,
    ! [mode 3: ninvar and the invar values, unused on this path], psi
    character(len=*), intent(in) :: fname
    real(8) :: xmaxp, ymaxp, zmaxp, hxp, hyp, hzp
    real(8), allocatable :: invar(:)
    integer :: nxp, nyp, nzp, ninvar
    open(1, file=fname, status='old')
    call titols(1)
    read(1, *) xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, nxp, nyp, nzp
    call check_grid(nxp, nyp, nzp)
    read(1, *) rimp
    read(1, *) vimp
    if (mode == 3) then
      read(1, *) ninvar
      allocate(invar(ninvar))
      read(1, *) invar
    end if
    read(1, *) psi
    close(1)
  end subroutine read_mode23

  subroutine read_mode0(fname)
    ! readenc mode 0 (readenc.f90:80-100): a density, or else a complex
    ! wave function, after a grid line that also carries limp and an
    ! impurity position (read and not used: the namelist's stays).  The
    ! reference tells them apart by the error a real read of "(re,im)" gives;
    ! that error depends on the compiler's runtime, so here the first value
    ! is looked at: a complex one opens with "(".  A density becomes
    ! psi = sqrt(den) exp(i k.r), the boost of DFT4He3d.f90:606-632.  The
    ! reference forms the phase with cmplx() of default kind, so its cosine
    ! and sine are rounded to REAL(4); the same rounding is kept here.
    character(len=*), intent(in) :: fname
    real(8) :: xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, xf, yf, zf, ph, xx, yy, zz
    real(8), allocatable :: den(:,:,:)
    integer :: nxp, nyp, nzp, ix, iy, iz
    logical :: limpf
    character(len=64) :: head
    allocate(den(nx,ny,nz))
    open(1, file=fname, status='old')
    call titols(1)
    read(1, *) xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, nxp, nyp, nzp, limpf, xf, yf, zf
    call check_grid(nxp, nyp, nzp)
    read(1, '(A)') head
    backspace(1)
    head = adjustl(head)
    if (head(1:1) == '(') then
      read(1, *) psi
      write(6, '("From Readenc: We have read a complex w.f.")')
    else
      read(1, *) den
      write(6, '("From Readenc: We have read a density")')
      if (xlamda /= 0.d0) xlamdaz = xlamda*ktops/(2.d0*h2o2m4)
      do iz = 1, nz
        zz = -zmax + hz*(iz - 1)
        do iy = 1, ny
          yy = -ymax + hy*(iy - 1)
          do ix = 1, nx
            xx = -xmax + hx*(ix - 1)
            ph = xx*xlamdax + yy*xlamday + zz*xlamdaz
            psi(ix, iy, iz) = sqrt(den(ix, iy, iz))*cmplx(real(cos(ph), 4), real(sin(ph), 4), kind=8)
          end do
        end do
      end do
    end if
    close(1)
  end subroutine read_mode0

  subroutine moments()
    ! centre of mass (r_cm.f90:23-55), <Lx,Ly,Lz> about it (DFT4He3d.f90:883-925)
    ! and <x2,y2,z2> (DFT4He3d.f90:935-947), reduced on the device
    real(8) :: mom(10)
    call hetd_check(hetd_moments(hctx, dble(n4), mom), 'moments')
    cm = mom(1:3)
    r2 = mom(4:6)
    ang = mom(8:10)
  end subroutine moments

  subroutine write_mode2(fname, it, ttime)
    ! printoutc layout (printoutc.f90:52-76)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: it
    real(8), intent(in) :: ttime
    call hetd_check(hetd_get_field(hctx, HETD_F_PSI, c_loc(psi)), 'get_field')
    call hetd_check(hetd_get_imp(hctx, rimp, vimp, aimp), 'get_imp')
    call energy()
    call hetd_check(hetd_norm(hctx, auxn4), 'norm')
    call moments()
    open(10, file=fname, form='FORMATTED')
    write(10, '("#  ")')
    write(10, '("#  Density after ",I15," iterations")') it
    write(10, '("#  Actual deltatps....:",1p,E20.10)') deltatps
    write(10, '("#  Total evolution time(ps)...:",1p,E20.12)') ttime
    write(10, '("#  Impurity potential.........:  ",A)') selec_gs
    write(10, '("#  r_cutoff_gs:",1p,E16.8)') r_cutoff_gs
    write(10, '("#  umax_gs....:",1p,E16.8)') umax_gs
    write(10, '("#  xcm,ycm,zcm:",1p,3E16.8)') cm
    write(10, '("#  NParticles.:",1p,E16.8)') auxn4
    write(10, '("#  Ekin.......:",1p,E16.8)') ekin4
    write(10, '("#  Elj........:",1p,E16.8)') elj4
    write(10, '("#  Ealphas....:",1p,E16.8)') ealphas
    write(10, '("#  Ecorr......:",1p,E16.8)') ecor4
    write(10, '("#  Esolid.....:",1p,E16.8)') esolid
    write(10, '("#  Ekinx......:",1p,E16.8)') ekinx
    write(10, '("#  Eimpu......:",1p,E16.8)') eimpu
    write(10, '("#  Etot.......:",1p,E16.8)') etot
    write(10, '("#  <Lx,y,z>...:",1p,3E16.8)') ang
    write(10, '("#  <x2,y2,z2>.:",1p,3E16.8)') r2
    write(10, '("#  ")')
    write(10, *) xmax, ymax, zmax, hx, hy, hz, nx, ny, nz
    write(10, *) rimp
    write(10, *) vimp
    write(10, '(2("(",ES24.16E3,",",ES24.16E3,")",1x))') psi
    close(10)
    if (lhistory) call write_history(trim(fname)//'.hist')
  end subroutine write_mode2

  subroutine write_history(fname)
    character(len=*), intent(in) :: fname
    integer(c_int8_t), allocatable, target :: blob(:)
    allocate(blob(hetd_history_size(hctx)))
    call hetd_check(hetd_save_history(hctx, c_loc(blob)), 'save_history')
    open(11, file=fname, access='stream', form='unformatted', status='replace')
    write(11) blob
    close(11)
  end subroutine write_history

  subroutine read_history(fname)
    character(len=*), intent(in) :: fname
    integer(c_int8_t), allocatable, target :: blob(:)
    integer(8) :: nbytes
    logical :: there
    inquire(file=fname, exist=there, size=nbytes)
    if (.not. there) then
      write(6, '(" No integrator history in ",A,": the run starts with Runge-Kutta steps")') fname
      return
    end if
    allocate(blob(nbytes))
    open(11, file=fname, access='stream', form='unformatted', status='old')
    read(11) blob
    close(11)
    call hetd_check(hetd_load_history(hctx, c_loc(blob), int(nbytes, c_int64_t)), 'load_history')
    call hetd_check(hetd_get_imp(hctx, rimp, vimp, aimp), 'get_imp')
    write(6, '(" Integrator history restored from ",A)') fname
  end subroutine read_history

end program hetd_drive
