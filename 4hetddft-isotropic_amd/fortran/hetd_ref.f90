!-----------------------------------------------------------------------------
! hetd_ref -- the reference's call surface for the propagation step, backed by
! the gfx950 library through hetd_capi.  Same routine names and argument
! lists as the reference; the state they act on lives in device memory
! (one context per run, module variable hctx) instead of module arrays.
!   fftini   fft.f90:5-76          potenimp  potenimp.f90:1-20
!   poten    poten.f90:5-118       forceimp  potenimp.f90:24-55
!   steprk   steprk.f90:1-153      steppc    steppc.f90:1-178
!   energy   energy.f90:12-144     (energies as in module energies,
!                                   modules.f90:41-52)
!   derivnD  init_deriv_parallel.f90:2-48 (the 3-D real and complex forms)
! A failing call STOPs with the library's message, as the reference STOPs.
!-----------------------------------------------------------------------------
module hetd_ref
  use hetd_capi
  implicit none
  type(c_ptr), save :: hctx = c_null_ptr
  real(8), save :: etot = 0, etot4 = 0, ekin4 = 0, ekinx = 0, elj4 = 0, ealphas = 0
  real(8), save :: ecor4 = 0, eimpu = 0, eHeX = 0, esolid = 0
  ! grid, spacing and boundary rule of the context, for derivnD's checks
  integer, save :: hgrid(3) = 0, hicon = 0
  real(8), save :: hspace(3) = 0

  ! derivnD (init_deriv_parallel.f90:2-48): the id-th derivative along
  ! axis iv with the context's npd-point stencil
  interface derivnD
    module procedure deriv3d_p, deriv3d_pc
  end interface derivnD

contains

  subroutine hetd_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= HETD_OK) then
      write(0, '(A,": ",A)') what, trim(hetd_message(hctx))
      if (rc == HETD_ERR_GUARD) error stop 'den not > 0 (steprk)'
      error stop 'hetd: device call failed'
    end if
  end subroutine hetd_check

  ! The FFT plans, stencil weights and form factors are made by hetd_init
  ! (hetd_setup); fftini remains for the call surface.
  subroutine fftini(nx, ny, nz, nthread)
    integer(4), intent(in) :: nx, ny, nz, nthread
    if (.not. c_associated(hctx)) error stop 'fftini: hetd_setup has not been called'
    if (nx < 1 .or. ny < 1 .or. nz < 1 .or. nthread < 0) error stop 'fftini: bad arguments'
  end subroutine fftini

  subroutine hetd_setup(p, selec, r_cutoff, umax)
    type(hetd_params), intent(in) :: p
    character(len=*), intent(in) :: selec
    real(8), intent(in) :: r_cutoff, umax
    real(8), allocatable, target :: potion(:)
    integer(c_int32_t) :: npot
    real(c_double) :: delinter
    integer(c_int) :: rc
    character(kind=c_char, len=81) :: cs
    cs = trim(selec)//c_null_char
    ! potenimpini (potenimpini.f90:11-25); the table is built on the host, so
    ! an unknown selec stops before the device is touched
    rc = hetd_build_potion(cs, r_cutoff, umax, max(p%nx, p%ny, p%nz), max(p%xmax, p%ymax, p%zmax), &
                           c_null_ptr, npot, delinter)
    if (rc /= HETD_OK) then
      write(0, '(A,": aquest potencial no esta definit")') trim(selec)
      error stop 'From V_impur: 0001'
    end if
    allocate(potion(npot))
    rc = hetd_build_potion(cs, r_cutoff, umax, max(p%nx, p%ny, p%nz), max(p%xmax, p%ymax, p%zmax), &
                           c_loc(potion), npot, delinter)
    call hetd_check(rc, 'hetd_build_potion')
    rc = hetd_init(p, hctx)
    call hetd_check(rc, 'hetd_init')
    hgrid = [p%nx, p%ny, p%nz]
    hicon = p%icon
    hspace = 2.0d0*abs([p%xmax, p%ymax, p%zmax])/hgrid  ! DFT4He3d.f90:274
    rc = hetd_set_potion(hctx, potion, npot, delinter)
    call hetd_check(rc, 'hetd_set_potion')
  end subroutine hetd_setup

  subroutine potenimp(rimp)
    real(8), intent(in) :: rimp(3)
    call hetd_check(hetd_potenimp(hctx, rimp), 'potenimp')
  end subroutine potenimp

  subroutine poten()
    call hetd_check(hetd_poten(hctx), 'poten')
  end subroutine poten

  subroutine forceimp(rimp, F)
    real(8), intent(in) :: rimp(3)
    real(8), intent(out) :: F(3)
    call hetd_check(hetd_forceimp(hctx, rimp, F), 'forceimp')
  end subroutine forceimp

  subroutine steprk(deltat)
    real(8), intent(in) :: deltat
    call hetd_check(hetd_steprk(hctx, deltat), 'steprk')
  end subroutine steprk

  subroutine steppc(deltat, errHe, errimp)
    real(8), intent(in) :: deltat
    real(8), intent(out) :: errHe, errimp
    call hetd_check(hetd_steppc(hctx, deltat, errHe, errimp), 'steppc')
  end subroutine steppc

  subroutine energy()
    type(hetd_energies) :: e
    call hetd_check(hetd_energy(hctx, e), 'energy')
    etot = e%etot; etot4 = e%etot4; ekin4 = e%ekin4; ekinx = e%ekinx; elj4 = e%elj4
    ealphas = e%ealphas; ecor4 = e%ecor4; eimpu = e%eimpu; eHeX = e%eHeX; esolid = e%esolid
  end subroutine energy

  ! The library differentiates with the stencil, spacing and boundary rule
  ! of its context: a call for another grid, spacing or icon stops.
  subroutine deriv_args(id, nn, h, iv, icon)
    integer(4), intent(in) :: id, nn(3), iv, icon
    real(8), intent(in) :: h
    if (.not. c_associated(hctx)) error stop 'derivnD: hetd_setup has not been called'
    if (id < 1 .or. id > 2 .or. iv < 1 .or. iv > 3) error stop 'derivnD: id must be 1 or 2, iv 1 to 3'
    if (any(nn /= hgrid)) error stop 'derivnD: the grid differs from the context''s'
    if (icon /= hicon) error stop 'derivnD: icon differs from the context''s'
    if (abs(h - hspace(iv)) > 1.0d-12*hspace(iv)) error stop 'derivnD: h differs from the context''s spacing'
  end subroutine deriv_args

  subroutine deriv3d_p(id, nn, h, iv, f, df, icon)
    integer(4), intent(in) :: id, nn(3), iv, icon
    real(8), intent(in) :: h
    real(8), intent(in), target :: f(nn(1), nn(2), nn(3))
    real(8), intent(out), target :: df(nn(1), nn(2), nn(3))
    call deriv_args(id, nn, h, iv, icon)
    call hetd_check(hetd_deriv(hctx, id, iv, 0, c_loc(f), c_loc(df)), 'derivnD')
  end subroutine deriv3d_p

  subroutine deriv3d_pc(id, nn, h, iv, f, df, icon)
    integer(4), intent(in) :: id, nn(3), iv, icon
    real(8), intent(in) :: h
    complex(8), intent(in), target :: f(nn(1), nn(2), nn(3))
    complex(8), intent(out), target :: df(nn(1), nn(2), nn(3))
    call deriv_args(id, nn, h, iv, icon)
    call hetd_check(hetd_deriv(hctx, id, iv, 1, c_loc(f), c_loc(df)), 'derivnD')
  end subroutine deriv3d_pc

end module hetd_ref
