! ISO_C_BINDING layer over libmoloch_hip (include/moloch_hip.h).
!
! This is the Fortran side of the drop-in for Main/mod_moloch.F90:127
!   public :: allocate_moloch, init_moloch, moloch
! It keeps those three phases under the same names with a _gpu suffix and
! moves the mo_atm arrays across the C-ABI with their own Fortran bounds
! (Main/mod_atm_interface.F90:579-624), so the host layout stays canonical
! and the device keeps its padded layout.  Every non-zero status from the
! library stops the run with the library's message, as fatal() does in the
! reference (Main/abort.F90:17-33).
module mod_moloch_gpu

  use, intrinsic :: iso_c_binding

  implicit none

  private

  integer(c_int), parameter, public :: mlc_ok = 0
  integer(c_int), parameter, public :: mlc_e_nonfinite = -4
  integer(c_int), parameter, public :: mlc_fp64 = 0, mlc_mixed = 1
  integer(c_int), parameter, public :: mlc_param_tkemin = 1, mlc_param_mo_h = 2

  ! field identifiers (enum mlc_field)
  integer(c_int), parameter, public :: &
    mlc_u = 1, mlc_v = 2, mlc_w = 3, mlc_pai = 4, mlc_tetav = 5, mlc_t = 6, &
    mlc_tvirt = 7, mlc_qx = 8, mlc_ux = 9, mlc_vx = 10, mlc_p = 11,         &
    mlc_rho = 12, mlc_qsat = 13, mlc_ps = 14, mlc_tten = 15, mlc_uten = 16, &
    mlc_vten = 17, mlc_qxten = 18,                                          &
    mlc_tke = 30, mlc_tketen = 31, mlc_trac = 32, mlc_chiten = 33,          &
    mlc_fmz = 100, mlc_fmzf = 101, mlc_rfmzu = 102, mlc_rfmzv = 103,        &
    mlc_bdywtu = 104, mlc_bdywtv = 105, mlc_bdywtw = 106, mlc_zeta = 107,   &
    mlc_dz = 108, mlc_mx = 109, mlc_mu = 110, mlc_mv = 111, mlc_hx = 112,   &
    mlc_hy = 113, mlc_coru = 114, mlc_corv = 115, mlc_rlat = 116,           &
    mlc_ffilt = 117
  ! surface geopotential of the terrain metrics (enum mlc_terrain_field):
  ! mddom%ht, htu, htv, with mlc_param_mo_h
  integer(c_int), parameter, public :: mlc_ht = 122, mlc_htu = 123, mlc_htv = 124
  ! launch-shape variant: the stages that form the terrain metrics
  integer(c_int), parameter, public :: mlc_var_terrain = 11
  integer(c_int), parameter, public :: mlc_var_mom_chunk = 12
  integer(c_int), parameter, public :: mlc_var_stagger_chunk = 13

  ! lateral-boundary inputs (enum mlc_lbc_field): b0/b1 of the v3dbound /
  ! v2dbound arrays, the Davies weights of ba_ud, ba_vd, ba_cr and tnudge
  integer(c_int), parameter, public :: &
    mlc_bnd_ps0 = 200, mlc_bnd_ps1 = 201, mlc_bnd_u0 = 202, mlc_bnd_u1 = 203, &
    mlc_bnd_v0 = 204, mlc_bnd_v1 = 205, mlc_bnd_t0 = 206, mlc_bnd_t1 = 207,   &
    mlc_bnd_pai0 = 208, mlc_bnd_pai1 = 209, mlc_bnd_qv0 = 210,                &
    mlc_bnd_qv1 = 211, mlc_bnd_qc0 = 212, mlc_bnd_qc1 = 213,                  &
    mlc_bnd_qi0 = 214, mlc_bnd_qi1 = 215, mlc_relax_u = 216,                  &
    mlc_relax_v = 217, mlc_relax_c = 218, mlc_tnudge = 219

  ! mkslice inputs and outputs (enum mlc_slice_field): mo_atm%zetaf, xlat
  ! and the atms slices pf3d, th3d, rhb3d, wpx3d, rhox2d, tp2d, th700,
  ! ptrop, ktrop and kmxpbl (the level numbers as reals)
  integer(c_int), parameter, public :: &
    mlc_zetaf = 120, mlc_xlat = 121, mlc_sl_pf = 300, mlc_sl_th = 301,      &
    mlc_sl_rh = 302, mlc_sl_omega = 303, mlc_sl_rhox = 304, mlc_sl_tp = 305, &
    mlc_sl_th700 = 306, mlc_sl_ptrop = 307, mlc_sl_ktrop = 308,             &
    mlc_sl_kmxpbl = 309

  type, bind(c), public :: mlc_slice_desc
    real(c_double) :: rhmin, rhmax, calday, dayspy
    integer(c_int) :: icldmstrat, irceideal, ibltyp, reserved
  end type mlc_slice_desc

  ! horizontal diffusion of the MM cores (enum mlc_diff_field, diffu_x forms)
  integer(c_int), parameter, public :: &
    mlc_df_ht = 400, mlc_df_msfd = 401, mlc_df_pc = 402, mlc_df_pd = 403,   &
    mlc_df_ud = 404, mlc_df_vd = 405, mlc_df_wx = 406, mlc_df_uten = 407,   &
    mlc_df_vten = 408, mlc_df_f = 409, mlc_df_ften = 410,                   &
    mlc_df_hgfact = 411, mlc_df_xkc = 412, mlc_df_xkd = 413,                &
    mlc_df_xkcf = 414
  integer(c_int), parameter, public :: &
    mlc_diffu_x3d = 0, mlc_diffu_x3df = 1, mlc_diffu_x4d = 2

  type, bind(c), public :: mlc_diff_desc
    integer(c_int) :: jx, iy, kz, idiffu, idynamic, diffu_hgtf, nspec, reserved
    real(c_double) :: dx, dtsec, ckh, adyndif
  end type mlc_diff_desc

  type, bind(c), public :: mlc_lbc_desc
    integer(c_int) :: present_qc, present_qi, top_nudge, nztop
    integer(c_int) :: spectral_nudge, km, lm, reserved
    real(c_double) :: dtrad
  end type mlc_lbc_desc

  type, bind(c), public :: mlc_grid_desc
    integer(c_int) :: jx, iy, kz
    integer(c_int) :: nqx, ipptls, nadv, nsound
    integer(c_int) :: do_divdamp, do_divfilter, lrotllr, precision
    integer(c_int) :: ibltyp, ntr
    integer(c_int) :: crm
    integer(c_int) :: reserved(2)
    real(c_double) :: dx, dtsec, mo_ztop, mo_a0
  end type mlc_grid_desc

  interface
    integer(c_int) function mlc_create(g, h) bind(c, name='mlc_create')
      import :: c_int, c_ptr, mlc_grid_desc
      type(mlc_grid_desc), intent(in) :: g
      type(c_ptr), intent(out) :: h
    end function mlc_create
    integer(c_int) function mlc_destroy(h) bind(c, name='mlc_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_destroy
    integer(c_int) function mlc_set_static(h, f, a, lb, ub) &
        bind(c, name='mlc_set_static')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: f
      real(c_double), intent(in) :: a(*)
      integer(c_int), intent(in) :: lb(4), ub(4)
    end function mlc_set_static
    integer(c_int) function mlc_put_state(h, f, a, lb, ub) &
        bind(c, name='mlc_put_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: f
      real(c_double), intent(in) :: a(*)
      integer(c_int), intent(in) :: lb(4), ub(4)
    end function mlc_put_state
    integer(c_int) function mlc_get_state(h, f, a, lb, ub) &
        bind(c, name='mlc_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: f
      real(c_double), intent(inout) :: a(*)
      integer(c_int), intent(in) :: lb(4), ub(4)
    end function mlc_get_state
    integer(c_int) function mlc_init(h) bind(c, name='mlc_init')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_init
    integer(c_int) function mlc_step(h, n) bind(c, name='mlc_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n
    end function mlc_step
    integer(c_int) function mlc_reset_tendencies(h) &
        bind(c, name='mlc_reset_tendencies')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_reset_tendencies
    integer(c_int) function mlc_dynamical_core(h) &
        bind(c, name='mlc_dynamical_core')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_dynamical_core
    integer(c_int) function mlc_diagnostics(h) bind(c, name='mlc_diagnostics')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diagnostics
    integer(c_int) function mlc_status_update(h, dtinc) &
        bind(c, name='mlc_status_update')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dtinc
    end function mlc_status_update
    integer(c_int) function mlc_get_variant(h, stage) &
        bind(c, name='mlc_get_variant')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: stage
    end function mlc_get_variant
    integer(c_int) function mlc_set_param(h, param, value) &
        bind(c, name='mlc_set_param')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: param
      real(c_double), value :: value
    end function mlc_set_param
    integer(c_int) function mlc_massck(h, out) bind(c, name='mlc_massck')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(4)
    end function mlc_massck
    integer(c_int) function mlc_check_finite(h, pmax, pmin) &
        bind(c, name='mlc_check_finite')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: pmax, pmin
    end function mlc_check_finite
    integer(c_int) function mlc_synchronize(h) bind(c, name='mlc_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_synchronize
    integer(c_int) function mlc_lbc_configure(h, d) bind(c, name='mlc_lbc_configure')
      import :: c_int, c_ptr, mlc_lbc_desc
      type(c_ptr), value :: h
      type(mlc_lbc_desc), intent(in) :: d
    end function mlc_lbc_configure
    integer(c_int) function mlc_lbc_spectral(h, bvx, bvy, cnudge) &
        bind(c, name='mlc_lbc_spectral')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: bvx(*), bvy(*), cnudge(*)
    end function mlc_lbc_spectral
    integer(c_int) function mlc_boundary(h, x1) bind(c, name='mlc_boundary')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: x1
    end function mlc_boundary
    integer(c_int) function mlc_step_lbc(h, n, x1) bind(c, name='mlc_step_lbc')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_double), intent(in) :: x1(*)
    end function mlc_step_lbc
    integer(c_int) function mlc_diff_create(d, h) bind(c, name='mlc_diff_create')
      import :: c_int, c_ptr, mlc_diff_desc
      type(mlc_diff_desc), intent(in) :: d
      type(c_ptr), intent(out) :: h
    end function mlc_diff_create
    integer(c_int) function mlc_diff_destroy(h) bind(c, name='mlc_diff_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diff_destroy
    integer(c_int) function mlc_diff_put(h, f, a, lb, ub) bind(c, name='mlc_diff_put')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: f
      real(c_double), intent(in) :: a(*)
      integer(c_int), intent(in) :: lb(4), ub(4)
    end function mlc_diff_put
    integer(c_int) function mlc_diff_get(h, f, a, lb, ub) bind(c, name='mlc_diff_get')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: f
      real(c_double), intent(out) :: a(*)
      integer(c_int), intent(in) :: lb(4), ub(4)
    end function mlc_diff_get
    integer(c_int) function mlc_diff_initialize(h) bind(c, name='mlc_diff_initialize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diff_initialize
    integer(c_int) function mlc_diff_calc_coeff(h) bind(c, name='mlc_diff_calc_coeff')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diff_calc_coeff
    integer(c_int) function mlc_diff_d(h) bind(c, name='mlc_diff_d')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diff_d
    integer(c_int) function mlc_diff_x(h, form, n1, n2, fac) bind(c, name='mlc_diff_x')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: form, n1, n2
      real(c_double), value :: fac
    end function mlc_diff_x
    integer(c_int) function mlc_diff_synchronize(h) bind(c, name='mlc_diff_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function mlc_diff_synchronize
    integer(c_int) function mlc_mkslice(h, d) bind(c, name='mlc_mkslice')
      import :: c_int, c_ptr, mlc_slice_desc
      type(c_ptr), value :: h
      type(mlc_slice_desc), intent(in) :: d
    end function mlc_mkslice
    type(c_ptr) function mlc_last_error() bind(c, name='mlc_last_error')
      import :: c_ptr
    end function mlc_last_error
    integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

  public :: mlc_create, mlc_destroy, mlc_set_static, mlc_put_state
  public :: mlc_get_state, mlc_init, mlc_step, mlc_reset_tendencies
  public :: mlc_dynamical_core, mlc_diagnostics, mlc_status_update, mlc_set_param
  public :: mlc_get_variant
  public :: mlc_massck, mlc_check_finite, mlc_synchronize
  public :: mlc_lbc_configure, mlc_lbc_spectral, mlc_boundary, mlc_step_lbc
  public :: mlc_mkslice, mkslice_gpu
  public :: mlc_diff_create, mlc_diff_destroy, mlc_diff_put, mlc_diff_get
  public :: mlc_diff_initialize, mlc_diff_calc_coeff, mlc_diff_d, mlc_diff_x
  public :: mlc_diff_synchronize

  ! the drop-in phases and their transfer helpers
  public :: lbc_configure_gpu, lbc_spectral_gpu, lbc_relax_gpu, boundary_gpu, moloch_lbc_gpu
  public :: allocate_moloch_gpu, init_moloch_gpu, moloch_gpu, destroy_moloch_gpu
  public :: gpu_put, gpu_get, gpu_static, gpu_check

  type(c_ptr), public, save :: mlc_h = c_null_ptr

  interface gpu_put
    module procedure put_2d
    module procedure put_3d
    module procedure put_4d
    module procedure put_1d
  end interface gpu_put

  interface gpu_get
    module procedure get_2d
    module procedure get_3d
    module procedure get_4d
  end interface gpu_get

  interface gpu_static
    module procedure static_1d
    module procedure static_2d
    module procedure static_3d
  end interface gpu_static

  contains

  ! Stop with the library's message on a non-zero status.
  subroutine gpu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if ( rc == mlc_ok ) return
    p = mlc_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(0,'(a,a,i0,a,*(a))') what, ' failed, status ', rc, ': ', msg(1:n)
    error stop 'moloch_gpu: fatal'
  end subroutine gpu_check

  ! allocate_moloch (Main/mod_moloch.F90:159-199)
  subroutine allocate_moloch_gpu(g)
    type(mlc_grid_desc), intent(in) :: g
    call gpu_check(mlc_create(g, mlc_h), 'mlc_create')
  end subroutine allocate_moloch_gpu

  ! init_moloch (Main/mod_moloch.F90:201-308); statics must be set first
  subroutine init_moloch_gpu
    call gpu_check(mlc_init(mlc_h), 'mlc_init')
  end subroutine init_moloch_gpu

  ! moloch (Main/mod_moloch.F90:312-446) with physics off: nsteps steps
  subroutine moloch_gpu(nsteps)
    integer, intent(in) :: nsteps
    call gpu_check(mlc_step(mlc_h, int(nsteps, c_int)), 'mlc_step')
  end subroutine moloch_gpu

  ! do_apply_bdy (Main/mod_moloch.F90:305): enable the device boundary stage
  subroutine lbc_configure_gpu(present_qc, present_qi, top_nudge, nztop)
    logical, intent(in) :: present_qc, present_qi, top_nudge
    integer, intent(in) :: nztop
    type(mlc_lbc_desc) :: d
    d%present_qc = merge(1, 0, present_qc)
    d%present_qi = merge(1, 0, present_qi)
    d%top_nudge = merge(1, 0, top_nudge)
    d%nztop = int(nztop, c_int)
    d%spectral_nudge = 0
    d%km = 0
    d%lm = 0
    d%reserved = 0
    d%dtrad = 0.0_c_double
    call gpu_check(mlc_lbc_configure(mlc_h, d), 'mlc_lbc_configure')
  end subroutine lbc_configure_gpu

  ! mo_spectral_nudge: the bases bvx(jde1:jde2,1:2km), bvy(ide1:ide2,1:2lm)
  ! and cnudge(1:kz) of lowpass_init (Main/mod_bdycod.F90:3844-3896)
  subroutine lbc_spectral_gpu(present_qc, present_qi, top_nudge, nztop, km, lm, dtrad, &
                              bvx, bvy, cnudge)
    logical, intent(in) :: present_qc, present_qi, top_nudge
    integer, intent(in) :: nztop, km, lm
    real(c_double), intent(in) :: dtrad
    real(c_double), intent(in) :: bvx(:,:), bvy(:,:), cnudge(:)
    type(mlc_lbc_desc) :: d
    d%present_qc = merge(1, 0, present_qc)
    d%present_qi = merge(1, 0, present_qi)
    d%top_nudge = merge(1, 0, top_nudge)
    d%nztop = int(nztop, c_int)
    d%spectral_nudge = 1
    d%km = int(km, c_int)
    d%lm = int(lm, c_int)
    d%reserved = 0
    d%dtrad = dtrad
    call gpu_check(mlc_lbc_configure(mlc_h, d), 'mlc_lbc_configure')
    call gpu_check(mlc_lbc_spectral(mlc_h, bvx, bvy, cnudge), 'mlc_lbc_spectral')
  end subroutine lbc_spectral_gpu

  ! Davies weights of one boundary area from RegCM's own ring index
  ! ba%ibnd(jde1:jde2,ide1:ide2) and hefc(1:nspgx,1:kz)
  ! (Main/mod_bdycod.F90:3996-4031): hefc(ibnd(j,i),k) where ibnd > 0, -1
  ! (no relaxation) elsewhere
  subroutine lbc_relax_gpu(f, ibnd, hefc, kz)
    integer(c_int), intent(in) :: f
    integer, intent(in) :: ibnd(:,:)
    real(c_double), intent(in) :: hefc(:,:)
    integer, intent(in) :: kz
    real(c_double), pointer, contiguous :: xf(:,:,:)
    integer :: j, i, k
    allocate(xf(size(ibnd,1), size(ibnd,2), kz))
    do k = 1, kz
      do i = 1, size(ibnd,2)
        do j = 1, size(ibnd,1)
          if ( ibnd(j,i) > 0 ) then
            xf(j,i,k) = hefc(ibnd(j,i),k)
          else
            xf(j,i,k) = -1.0_c_double
          end if
        end do
      end do
    end do
    call put_3d(f, xf)
    deallocate(xf)
  end subroutine lbc_relax_gpu

  ! boundary() (Main/mod_moloch.F90:448-529) with time weight x1
  subroutine boundary_gpu(x1)
    real(c_double), intent(in) :: x1
    call gpu_check(mlc_boundary(mlc_h, x1), 'mlc_boundary')
  end subroutine boundary_gpu

  ! moloch() with do_apply_bdy, one step per entry of x1
  subroutine moloch_lbc_gpu(x1)
    real(c_double), intent(in) :: x1(:)
    call gpu_check(mlc_step_lbc(mlc_h, int(size(x1), c_int), x1), 'mlc_step_lbc')
  end subroutine moloch_lbc_gpu

  ! mkslice (Main/mod_slice.F90:112-382) on the device state; zetaf and
  ! (irceideal /= 1) xlat must have been set with gpu_static
  subroutine mkslice_gpu(rhmin, rhmax, calday, dayspy, icldmstrat, irceideal, ibltyp)
    real(c_double), intent(in) :: rhmin, rhmax, calday, dayspy
    integer, intent(in) :: icldmstrat, irceideal, ibltyp
    type(mlc_slice_desc) :: d
    d%rhmin = rhmin
    d%rhmax = rhmax
    d%calday = calday
    d%dayspy = dayspy
    d%icldmstrat = icldmstrat
    d%irceideal = irceideal
    d%ibltyp = ibltyp
    d%reserved = 0
    call gpu_check(mlc_mkslice(mlc_h, d), 'mlc_mkslice')
  end subroutine mkslice_gpu

  subroutine destroy_moloch_gpu
    call gpu_check(mlc_destroy(mlc_h), 'mlc_destroy')
    mlc_h = c_null_ptr
  end subroutine destroy_moloch_gpu

  ! --- transfers of mo_atm-style pointer arrays with their own bounds ----
  ! a 1-D k profile (tnudge)
  subroutine put_1d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ 1, 1, lbound(a,1), 1 ]
    ub = [ 1, 1, ubound(a,1), 1 ]
    call gpu_check(mlc_put_state(mlc_h, f, a, lb, ub), 'mlc_put_state')
  end subroutine put_1d

  subroutine put_2d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), 1, 1 ]
    ub = [ ubound(a,1), ubound(a,2), 1, 1 ]
    call gpu_check(mlc_put_state(mlc_h, f, a, lb, ub), 'mlc_put_state')
  end subroutine put_2d

  subroutine put_3d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:,:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), lbound(a,3), 1 ]
    ub = [ ubound(a,1), ubound(a,2), ubound(a,3), 1 ]
    call gpu_check(mlc_put_state(mlc_h, f, a, lb, ub), 'mlc_put_state')
  end subroutine put_3d

  subroutine put_4d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:,:,:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = lbound(a)
    ub = ubound(a)
    call gpu_check(mlc_put_state(mlc_h, f, a, lb, ub), 'mlc_put_state')
  end subroutine put_4d

  subroutine get_2d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(inout) :: a(:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), 1, 1 ]
    ub = [ ubound(a,1), ubound(a,2), 1, 1 ]
    call gpu_check(mlc_get_state(mlc_h, f, a, lb, ub), 'mlc_get_state')
  end subroutine get_2d

  subroutine get_3d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(inout) :: a(:,:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), lbound(a,3), 1 ]
    ub = [ ubound(a,1), ubound(a,2), ubound(a,3), 1 ]
    call gpu_check(mlc_get_state(mlc_h, f, a, lb, ub), 'mlc_get_state')
  end subroutine get_3d

  subroutine get_4d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(inout) :: a(:,:,:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = lbound(a)
    ub = ubound(a)
    call gpu_check(mlc_get_state(mlc_h, f, a, lb, ub), 'mlc_get_state')
  end subroutine get_4d

  subroutine static_1d(f, a, along)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:)
    integer, intent(in) :: along          ! 2: an i profile (rlat), 3: a k profile (ffilt)
    integer(c_int) :: lb(4), ub(4)
    lb = 1
    ub = 1
    lb(along) = lbound(a,1)
    ub(along) = ubound(a,1)
    call gpu_check(mlc_set_static(mlc_h, f, a, lb, ub), 'mlc_set_static')
  end subroutine static_1d

  subroutine static_2d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), 1, 1 ]
    ub = [ ubound(a,1), ubound(a,2), 1, 1 ]
    call gpu_check(mlc_set_static(mlc_h, f, a, lb, ub), 'mlc_set_static')
  end subroutine static_2d

  subroutine static_3d(f, a)
    integer(c_int), intent(in) :: f
    real(c_double), pointer, contiguous, intent(in) :: a(:,:,:)
    integer(c_int) :: lb(4), ub(4)
    lb = [ lbound(a,1), lbound(a,2), lbound(a,3), 1 ]
    ub = [ ubound(a,1), ubound(a,2), ubound(a,3), 1 ]
    call gpu_check(mlc_set_static(mlc_h, f, a, lb, ub), 'mlc_set_static')
  end subroutine static_3d

end module mod_moloch_gpu
