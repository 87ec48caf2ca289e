! Fortran host driver of the MM-core horizontal-diffusion operator on the
! device (mod_moloch_gpu's mlc_diff_* bindings).  It mirrors the sequence of
! Main/mod_tendency.F90: calc_coeff (:260) then diffusion (:1434-1466):
! diffu_d on the dot-point winds, diffu_x3d on a cross field, diffu_x3df on
! a kz+1-level field and diffu_x4d on a range of species.
!
!   diffusion_gpu_driver <input> <output>
!
! input : int32 jx iy kz idiffu idynamic diffu_hgtf nspec
!         real64 dx dtsec ckh adyndif
!         real64 ht, msfd, psb, psdotb (jx,iy); ubd3d, vbd3d (jx,iy,kz);
!                wb3d (jx,iy,kz+1); uten, vten (jx,iy,kz);
!                f, ften (jx,iy,kz+1,nspec)
! output: real64 hgfact (jx,iy); xkc, xkd (jx,iy,kz); xkcf (jx,iy,kz+1);
!         uten, vten (jx,iy,kz); ften (jx,iy,kz+1,nspec) after
!         x3d on species 1, x3df (fac 0.7) on species 2 and x4d (fac 0.3,
!         levels 1..kz) on species 3..nspec
program diffusion_gpu_driver

  use, intrinsic :: iso_c_binding
  use mod_moloch_gpu

  implicit none

  integer(c_int) :: hdr(7)
  real(c_double) :: rh(4)
  type(mlc_diff_desc) :: dd
  type(c_ptr) :: h
  character(len=1024) :: fin, fout
  integer :: uin, uout, jx, iy, kz, ns
  real(c_double), allocatable :: ht(:,:), msfd(:,:), pc(:,:), pd(:,:), hg(:,:)
  real(c_double), allocatable :: ud(:,:,:), vd(:,:,:), wx(:,:,:), uten(:,:,:), &
      vten(:,:,:), xkc(:,:,:), xkd(:,:,:), xkcf(:,:,:)
  real(c_double), allocatable :: f(:,:,:,:), ften(:,:,:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', &
       status='old', action='read')
  read(uin) hdr, rh
  jx = hdr(1); iy = hdr(2); kz = hdr(3); ns = hdr(7)
  allocate(ht(jx,iy), msfd(jx,iy), pc(jx,iy), pd(jx,iy), hg(jx,iy))
  allocate(ud(jx,iy,kz), vd(jx,iy,kz), wx(jx,iy,kz+1), uten(jx,iy,kz), vten(jx,iy,kz))
  allocate(xkc(jx,iy,kz), xkd(jx,iy,kz), xkcf(jx,iy,kz+1))
  allocate(f(jx,iy,kz+1,ns), ften(jx,iy,kz+1,ns))
  read(uin) ht, msfd, pc, pd, ud, vd, wx, uten, vten, f, ften
  close(uin)

  dd%jx = jx; dd%iy = iy; dd%kz = kz
  dd%idiffu = hdr(4); dd%idynamic = hdr(5); dd%diffu_hgtf = hdr(6)
  dd%nspec = ns; dd%reserved = 0
  dd%dx = rh(1); dd%dtsec = rh(2); dd%ckh = rh(3); dd%adyndif = rh(4)
  call gpu_check(mlc_diff_create(dd, h), 'mlc_diff_create')

  call put(mlc_df_ht, ht, 1, 1)
  call put(mlc_df_msfd, msfd, 1, 1)
  call put(mlc_df_pc, pc, 1, 1)
  call put(mlc_df_pd, pd, 1, 1)
  call put(mlc_df_ud, ud, kz, 1)
  call put(mlc_df_vd, vd, kz, 1)
  call put(mlc_df_wx, wx, kz+1, 1)
  call put(mlc_df_uten, uten, kz, 1)
  call put(mlc_df_vten, vten, kz, 1)
  call put(mlc_df_f, f, kz+1, ns)
  call put(mlc_df_ften, ften, kz+1, ns)

  ! initialize_diffusion, then one MM step's calc_coeff and diffusion
  call gpu_check(mlc_diff_initialize(h), 'mlc_diff_initialize')
  call gpu_check(mlc_diff_calc_coeff(h), 'mlc_diff_calc_coeff')
  call gpu_check(mlc_diff_d(h), 'mlc_diff_d')
  call gpu_check(mlc_diff_x(h, mlc_diffu_x3d, 1, 1, 1.0_c_double), 'mlc_diff_x x3d')
  if ( ns >= 2 ) &
    call gpu_check(mlc_diff_x(h, mlc_diffu_x3df, 2, 2, 0.7_c_double), 'mlc_diff_x x3df')
  if ( ns >= 3 ) &
    call gpu_check(mlc_diff_x(h, mlc_diffu_x4d, 3, ns, 0.3_c_double), 'mlc_diff_x x4d')

  call get(mlc_df_hgfact, hg, 1, 1)
  call get(mlc_df_xkc, xkc, kz, 1)
  call get(mlc_df_xkd, xkd, kz, 1)
  call get(mlc_df_xkcf, xkcf, kz+1, 1)
  call get(mlc_df_uten, uten, kz, 1)
  call get(mlc_df_vten, vten, kz, 1)
  call get(mlc_df_ften, ften, kz+1, ns)
  call gpu_check(mlc_diff_destroy(h), 'mlc_diff_destroy')

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', &
       status='replace', action='write')
  write(uout) hg, xkc, xkd, xkcf, uten, vten, ften
  close(uout)

  contains

  ! whole-extent boxes: (1:jx, 1:iy, 1:nk, 1:nn)
  subroutine put(fid, a, nk, nn)
    integer(c_int), intent(in) :: fid
    real(c_double), intent(in) :: a(*)
    integer, intent(in) :: nk, nn
    integer(c_int) :: lb(4), ub(4)
    lb = [ 1, 1, 1, 1 ]
    ub = [ jx, iy, nk, nn ]
    call gpu_check(mlc_diff_put(h, fid, a, lb, ub), 'mlc_diff_put')
  end subroutine put

  subroutine get(fid, a, nk, nn)
    integer(c_int), intent(in) :: fid
    real(c_double), intent(out) :: a(*)
    integer, intent(in) :: nk, nn
    integer(c_int) :: lb(4), ub(4)
    lb = [ 1, 1, 1, 1 ]
    ub = [ jx, iy, nk, nn ]
    call gpu_check(mlc_diff_get(h, fid, a, lb, ub), 'mlc_diff_get')
  end subroutine get

end program diffusion_gpu_driver
