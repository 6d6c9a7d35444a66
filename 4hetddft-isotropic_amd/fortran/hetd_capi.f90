!-----------------------------------------------------------------------------
! hetd_capi -- ISO_C_BINDING interfaces to the C ABI of include/hetd.h.
! The derived types mirror the C structs field for field.
!-----------------------------------------------------------------------------
module hetd_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HETD_OK = 0, HETD_ERR_ARG = 1, HETD_ERR_HIP = 2, &
                               HETD_ERR_GUARD = 3, HETD_ERR_STATE = 4
  integer(c_int32_t), parameter :: HETD_FP64 = 0, HETD_MIXED = 1
  integer(c_int32_t), parameter :: HETD_CORE_OT = 0, HETD_CORE_OTC = 1
  integer(c_int32_t), parameter :: HETD_F_PSI = 0, HETD_F_DEN = 1, HETD_F_POT4 = 2
  integer(c_int32_t), parameter :: HETD_TRAJ_STRIDE = 11

  type, bind(C) :: hetd_params
    integer(c_int32_t) :: nx, ny, nz, npd, icon, core, lsolid, precision
    real(c_double) :: xmax, ymax, zmax
    real(c_double) :: eps4, sigma4, h4, cp4, cpp4, h2o2m4, alphas, l_alpha, den0s
    real(c_double) :: c_solid, beta_solid, den_m, mimpur
    real(c_double) :: Lambdah, txmean, txsurf, tymean, tysurf, tzmean, tzsurf
  end type hetd_params

  type, bind(C) :: hetd_energies
    real(c_double) :: etot4, ekin4, elj4, ecor4, esolid, ealphas
    real(c_double) :: ekinx, eHeX, eimpu, etot
  end type hetd_energies

  interface
    subroutine hetd_default_params(p) bind(C, name='hetd_default_params')
      import :: hetd_params
      type(hetd_params), intent(out) :: p
    end subroutine
    integer(c_int) function hetd_init(p, ctx) bind(C, name='hetd_init')
      import :: hetd_params, c_int, c_ptr
      type(hetd_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function hetd_finalize(ctx) bind(C, name='hetd_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function hetd_last_error(ctx) bind(C, name='hetd_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hetd_build_potion(selec, r_cutoff, umax, nmax, xmaxbox, potion, npot, delinter) &
        bind(C, name='hetd_build_potion')
      import :: c_int, c_int32_t, c_double, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: selec
      real(c_double), value :: r_cutoff, umax, xmaxbox
      integer(c_int32_t), value :: nmax
      type(c_ptr), value :: potion
      integer(c_int32_t), intent(out) :: npot
      real(c_double), intent(out) :: delinter
    end function
    integer(c_int) function hetd_set_potion(ctx, potion, npot, delinter) bind(C, name='hetd_set_potion')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: potion
      integer(c_int32_t), value :: npot
      real(c_double), value :: delinter
    end function
    integer(c_int) function hetd_set_state(ctx, psi, r, v) bind(C, name='hetd_set_state')
      import :: c_int, c_double, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), dimension(*), intent(in) :: psi
      real(c_double), dimension(3), intent(in) :: r, v
    end function
    integer(c_int) function hetd_potenimp(ctx, r) bind(C, name='hetd_potenimp')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), dimension(3), intent(in) :: r
    end function
    integer(c_int) function hetd_poten(ctx) bind(C, name='hetd_poten')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hetd_forceimp(ctx, r, F) bind(C, name='hetd_forceimp')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), dimension(3), intent(in) :: r
      real(c_double), dimension(3), intent(out) :: F
    end function
    integer(c_int) function hetd_set_aimp(ctx, a) bind(C, name='hetd_set_aimp')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), dimension(3), intent(in) :: a
    end function
    integer(c_int) function hetd_vsurf(lsurf3d, x, y, z, v) bind(C, name='hetd_vsurf')
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: lsurf3d
      real(c_double), value :: x, y, z
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function hetd_set_uext(ctx, uext) bind(C, name='hetd_set_uext')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, uext
    end function
    integer(c_int) function hetd_set_surface(ctx, lsurf3d, zsurf) bind(C, name='hetd_set_surface')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: lsurf3d
      real(c_double), value :: zsurf
    end function
    integer(c_int) function hetd_autotune(ctx, dt, steps, choice) bind(C, name='hetd_autotune')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int32_t), value :: steps
      integer(c_int32_t), intent(out) :: choice
    end function
    integer(c_int) function hetd_steprk(ctx, dt) bind(C, name='hetd_steprk')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function hetd_steppc(ctx, dt, errHe, errimp) bind(C, name='hetd_steppc')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), intent(out) :: errHe, errimp
    end function
    integer(c_int) function hetd_energy(ctx, e) bind(C, name='hetd_energy')
      import :: c_int, c_ptr, hetd_energies
      type(c_ptr), value :: ctx
      type(hetd_energies), intent(out) :: e
    end function
    integer(c_int) function hetd_norm(ctx, norm) bind(C, name='hetd_norm')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norm
    end function
    integer(c_int) function hetd_evolve(ctx, dt, nsteps, traj) bind(C, name='hetd_evolve')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int32_t), value :: nsteps
      real(c_double), dimension(*), intent(out) :: traj
    end function
    integer(c_int64_t) function hetd_history_size(ctx) bind(C, name='hetd_history_size')
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hetd_save_history(ctx, host) bind(C, name='hetd_save_history')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, host
    end function
    integer(c_int) function hetd_load_history(ctx, host, bytes) bind(C, name='hetd_load_history')
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, host
      integer(c_int64_t), value :: bytes
    end function
    integer(c_int) function hetd_moments(ctx, npart, out) bind(C, name='hetd_moments')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: npart
      real(c_double), dimension(10), intent(out) :: out
    end function
    integer(c_int) function hetd_pool_evolve(ctxs, n, dt, nsteps) bind(C, name='hetd_pool_evolve')
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), dimension(*), intent(in) :: ctxs
      integer(c_int32_t), value :: n, nsteps
      real(c_double), value :: dt
    end function
    integer(c_int) function hetd_get_field(ctx, which, host) bind(C, name='hetd_get_field')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      type(c_ptr), value :: host
    end function
    integer(c_int) function hetd_deriv(ctx, id, iv, is_complex, f, df) bind(C, name='hetd_deriv')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, f, df
      integer(c_int32_t), value :: id, iv, is_complex
    end function
    integer(c_int) function hetd_get_imp(ctx, r, v, a) bind(C, name='hetd_get_imp')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), dimension(3), intent(out) :: r, v, a
    end function
  end interface

contains

  ! the error text of the last failing call, as a Fortran string
  function hetd_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = hetd_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function hetd_message

end module hetd_capi
