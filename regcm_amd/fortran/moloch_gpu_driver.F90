! Fortran host driver of the GPU MOLOCH step.
!
! It holds the state the way RegCM does -- pointer arrays with the bounds of
! mo_atm / mddom on a single domain (Main/mod_atm_interface.F90:231-351,
! 579-624) -- and drives the step through mod_moloch_gpu: allocate, upload
! the static metrics and the state, init, step, download.  Input and output
! are raw stream files (native byte order) so that a harness can feed it
! synthetic domains and compare the result:
!
!   moloch_gpu_driver <input> <output>
!
! input : int32 jx iy kz nqx ipptls nadv nsound divdamp divfilter rotllr
!               nsteps mode (0: moloch steps, 1: the split physics path,
!                            2: moloch steps with lateral boundaries,
!                            3: moloch steps with TKE and tracers,
!                            4: moloch steps on a doubly periodic domain,
!                               i_crm = 1, every array over 1..jx x 1..iy)
!         real64 dx dtsec mo_ztop mo_a0 mo_h (mo_h <= 0: no orography given)
!         then the arrays listed in read_all, each over its own bounds
!         mode 2 then: int32 present_qc present_qi top_nudge nztop nspgx,
!         ibnd of ba_ud, ba_vd, ba_cr (jde x ide each), real64 hefc(nspgx,kz),
!         the boundary states b0/b1 of ps, u, v, t, pai, qv, qc, qi over the
!         bounds of the matching state arrays, tnudge(kz), x1(nsteps)
!         mode 3 then: int32 ntr, real64 tkemin, tke(jce x ice x kz+1),
!         trac(jce x ice x kz x ntr)
! output: the state arrays listed in write_all, then massck(4); mode 3
!         then tke and trac
program moloch_gpu_driver

  use, intrinsic :: iso_c_binding
  use mod_moloch_gpu

  implicit none

  integer(c_int) :: hdr(12)
  real(c_double) :: rh(5)
  integer :: jx, iy, kz, kzp1, nqx, nsteps, mode, n
  integer :: jde1, jde2, jdi1, jdi2, jce1, jce2, jci1, jci2
  integer :: ide1, ide2, idi1, idi2, ice1, ice2, ici1, ici2
  type(mlc_grid_desc) :: g
  character(len=1024) :: fin, fout
  integer :: uin, uout
  real(c_double) :: mass(4)

  ! mo_atm-like state
  real(c_double), pointer, contiguous, dimension(:,:,:) :: u, v, w, pai, tetav, &
      t, tvirt, ux, vx, p, rho, qsat, fmz, fmzf, rfmzu, rfmzv, bdywtu, bdywtv, &
      bdywtw, zeta, dz
  real(c_double), pointer, contiguous, dimension(:,:,:,:) :: qx
  real(c_double), pointer, contiguous, dimension(:,:) :: ps, mx, mu, mv, hx, hy, &
      coru, corv, ht, htu, htv
  real(c_double), pointer, contiguous, dimension(:) :: rlat, ffilt
  ! lateral boundaries (mode 2)
  integer(c_int) :: lhdr(5)
  integer, allocatable :: ib_ud(:,:), ib_vd(:,:), ib_cr(:,:)
  real(c_double), allocatable :: hefc(:,:), x1(:)
  real(c_double), pointer, contiguous, dimension(:,:,:) :: bu0, bu1, bv0, bv1, bt0, bt1, &
      bp0, bp1, bq0, bq1, bc0, bc1, bi0, bi1
  real(c_double), pointer, contiguous, dimension(:,:) :: bps0, bps1
  real(c_double), pointer, contiguous, dimension(:) :: tnudge
  ! TKE and tracers (mode 3): mo_atm%tke, mo_atm%trac
  integer(c_int) :: ntr
  real(c_double) :: tkemin
  real(c_double), pointer, contiguous, dimension(:,:,:) :: tke
  real(c_double), pointer, contiguous, dimension(:,:,:,:) :: trac

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', &
       status='old', action='read')
  read(uin) hdr, rh
  jx = hdr(1); iy = hdr(2); kz = hdr(3); nqx = hdr(4)
  nsteps = hdr(11); mode = hdr(12)
  kzp1 = kz + 1
  ! single-domain extents (Main/mod_atm_interface.F90:231-287)
  jde1 = 1; jde2 = jx; jdi1 = 2; jdi2 = jx-1; jce1 = 1; jce2 = jx-1
  jci1 = 2; jci2 = jx-2
  ide1 = 1; ide2 = iy; idi1 = 2; idi2 = iy-1; ice1 = 1; ice2 = iy-1
  ici1 = 2; ici2 = iy-2
  if ( mode == 4 ) then
    ! i_crm: the cross extents reach jx, iy and no edge is a boundary
    ! (Main/mpplib/mod_mppparam.F90:1290-1309)
    jdi1 = 1; jdi2 = jx; jce2 = jx; jci1 = 1; jci2 = jx
    idi1 = 1; idi2 = iy; ice2 = iy; ici1 = 1; ici2 = iy
  end if

  allocate(u(jde1:jde2,ice1:ice2,1:kz), v(jce1:jce2,ide1:ide2,1:kz))
  allocate(w(jce1:jce2,ice1:ice2,1:kzp1))
  allocate(pai(jce1:jce2,ice1:ice2,1:kz), tetav(jce1:jce2,ice1:ice2,1:kz))
  allocate(t(jce1:jce2,ice1:ice2,1:kz), tvirt(jce1:jce2,ice1:ice2,1:kz))
  allocate(qx(jce1:jce2,ice1:ice2,1:kz,1:nqx))
  allocate(ux(jce1:jce2,ice1:ice2,1:kz), vx(jce1:jce2,ice1:ice2,1:kz))
  allocate(p(jce1:jce2,ice1:ice2,1:kz), rho(jce1:jce2,ice1:ice2,1:kz))
  allocate(qsat(jce1:jce2,ice1:ice2,1:kz), ps(jce1:jce2,ice1:ice2))
  allocate(fmz(jce1:jce2,ice1:ice2,1:kz), fmzf(jce1:jce2,ice1:ice2,1:kzp1))
  allocate(rfmzu(jde1:jde2,ice1:ice2,1:kz), rfmzv(jce1:jce2,ide1:ide2,1:kz))
  allocate(bdywtu(jdi1:jdi2,ici1:ici2,1:kz), bdywtv(jci1:jci2,idi1:idi2,1:kz))
  allocate(bdywtw(jci1:jci2,ici1:ici2,1:kz))
  allocate(zeta(jce1:jce2,ice1:ice2,1:kz), dz(jce1:jce2,ice1:ice2,1:kz))
  allocate(mx(jde1:jde2,ide1:ide2), mu(jde1:jde2,ide1:ide2), mv(jde1:jde2,ide1:ide2))
  allocate(hx(jde1:jde2,ice1:ice2), hy(jce1:jce2,ide1:ide2))
  allocate(coru(jde1:jde2,ice1:ice2), corv(jce1:jce2,ide1:ide2))
  allocate(rlat(ide1:ide2), ffilt(1:kz))
  allocate(ht(jce1:jce2,ice1:ice2), htu(jde1:jde2,ice1:ice2), htv(jce1:jce2,ide1:ide2))
  call read_all
  if ( mode == 2 ) call read_lbc
  ntr = 0
  if ( mode == 3 ) then
    read(uin) ntr, tkemin
    allocate(tke(jce1:jce2,ice1:ice2,1:kzp1), trac(jce1:jce2,ice1:ice2,1:kz,1:ntr))
    read(uin) tke, trac
  end if
  close(uin)

  g%jx = jx; g%iy = iy; g%kz = kz; g%nqx = nqx
  g%ipptls = hdr(5); g%nadv = hdr(6); g%nsound = hdr(7)
  g%do_divdamp = hdr(8); g%do_divfilter = hdr(9); g%lrotllr = hdr(10)
  g%precision = mlc_fp64
  g%ibltyp = merge(2, 0, mode == 3); g%ntr = ntr; g%crm = merge(1, 0, mode == 4)
  g%reserved = 0
  g%dx = rh(1); g%dtsec = rh(2); g%mo_ztop = rh(3); g%mo_a0 = rh(4)

  call allocate_moloch_gpu(g)
  call gpu_static(mlc_fmz, fmz)
  call gpu_static(mlc_fmzf, fmzf)
  call gpu_static(mlc_rfmzu, rfmzu)
  call gpu_static(mlc_rfmzv, rfmzv)
  call gpu_static(mlc_bdywtu, bdywtu)
  call gpu_static(mlc_bdywtv, bdywtv)
  call gpu_static(mlc_bdywtw, bdywtw)
  call gpu_static(mlc_zeta, zeta)
  call gpu_static(mlc_dz, dz)
  call gpu_static(mlc_mx, mx)
  call gpu_static(mlc_mu, mu)
  call gpu_static(mlc_mv, mv)
  call gpu_static(mlc_hx, hx)
  call gpu_static(mlc_hy, hy)
  call gpu_static(mlc_coru, coru)
  call gpu_static(mlc_corv, corv)
  call gpu_static(mlc_rlat, rlat, 2)
  call gpu_static(mlc_ffilt, ffilt, 3)
  if ( rh(5) > 0.0_c_double ) then
    ! the orography the metrics were formed from (compute_moloch_static):
    ! the library forms them itself where they agree bit for bit
    call gpu_static(mlc_ht, ht)
    call gpu_static(mlc_htu, htu)
    call gpu_static(mlc_htv, htv)
    call gpu_check(mlc_set_param(mlc_h, mlc_param_mo_h, rh(5)), 'mlc_set_param')
  end if
  call gpu_put(mlc_u, u)
  call gpu_put(mlc_v, v)
  call gpu_put(mlc_w, w)
  call gpu_put(mlc_pai, pai)
  call gpu_put(mlc_tetav, tetav)
  call gpu_put(mlc_t, t)
  call gpu_put(mlc_tvirt, tvirt)
  call gpu_put(mlc_qx, qx)
  call gpu_put(mlc_ux, ux)
  call gpu_put(mlc_vx, vx)
  call gpu_put(mlc_p, p)
  call gpu_put(mlc_rho, rho)
  call gpu_put(mlc_qsat, qsat)
  call gpu_put(mlc_ps, ps)
  call init_moloch_gpu
  write(*,'(a,i0)') 'terrain metrics formed by stages ', mlc_get_variant(mlc_h, mlc_var_terrain)
  if ( mode == 3 ) then
    call gpu_check(mlc_set_param(mlc_h, mlc_param_tkemin, tkemin), 'mlc_set_param')
    call gpu_put(mlc_tke, tke)
    call gpu_put(mlc_trac, trac)
  end if

  if ( mode == 0 .or. mode == 3 .or. mode == 4 ) then
    call moloch_gpu(nsteps)
  else if ( mode == 2 ) then
    call lbc_configure_gpu(lhdr(1) /= 0, lhdr(2) /= 0, lhdr(3) /= 0, int(lhdr(4)))
    call gpu_put(mlc_bnd_ps0, bps0)
    call gpu_put(mlc_bnd_ps1, bps1)
    call gpu_put(mlc_bnd_u0, bu0)
    call gpu_put(mlc_bnd_u1, bu1)
    call gpu_put(mlc_bnd_v0, bv0)
    call gpu_put(mlc_bnd_v1, bv1)
    call gpu_put(mlc_bnd_t0, bt0)
    call gpu_put(mlc_bnd_t1, bt1)
    call gpu_put(mlc_bnd_pai0, bp0)
    call gpu_put(mlc_bnd_pai1, bp1)
    call gpu_put(mlc_bnd_qv0, bq0)
    call gpu_put(mlc_bnd_qv1, bq1)
    call gpu_put(mlc_bnd_qc0, bc0)
    call gpu_put(mlc_bnd_qc1, bc1)
    call gpu_put(mlc_bnd_qi0, bi0)
    call gpu_put(mlc_bnd_qi1, bi1)
    call lbc_relax_gpu(mlc_relax_u, ib_ud, hefc, kz)
    call lbc_relax_gpu(mlc_relax_v, ib_vd, hefc, kz)
    call lbc_relax_gpu(mlc_relax_c, ib_cr, hefc, kz)
    call gpu_put(mlc_tnudge, tnudge)
    call moloch_lbc_gpu(x1)
  else
    ! the sequence moloch() runs around a host physics package
    do n = 1, nsteps
      call gpu_check(mlc_reset_tendencies(mlc_h), 'mlc_reset_tendencies')
      call gpu_check(mlc_dynamical_core(mlc_h), 'mlc_dynamical_core')
      call gpu_check(mlc_diagnostics(mlc_h), 'mlc_diagnostics')
      call gpu_check(mlc_status_update(mlc_h, g%dtsec), 'mlc_status_update')
    end do
  end if

  call gpu_get(mlc_u, u)
  call gpu_get(mlc_v, v)
  call gpu_get(mlc_w, w)
  call gpu_get(mlc_pai, pai)
  call gpu_get(mlc_tetav, tetav)
  call gpu_get(mlc_t, t)
  call gpu_get(mlc_tvirt, tvirt)
  call gpu_get(mlc_qx, qx)
  call gpu_get(mlc_ux, ux)
  call gpu_get(mlc_vx, vx)
  call gpu_get(mlc_p, p)
  call gpu_get(mlc_rho, rho)
  call gpu_get(mlc_qsat, qsat)
  call gpu_get(mlc_ps, ps)
  if ( mode == 3 ) then
    call gpu_get(mlc_tke, tke)
    call gpu_get(mlc_trac, trac)
  end if
  call gpu_check(mlc_massck(mlc_h, mass), 'mlc_massck')
  call destroy_moloch_gpu

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', &
       status='replace', action='write')
  write(uout) u, v, w, pai, tetav, t, tvirt, qx, ux, vx, p, rho, qsat, ps, mass
  if ( mode == 3 ) write(uout) tke, trac
  close(uout)

  contains

  subroutine read_lbc
    integer :: nspgx
    read(uin) lhdr
    nspgx = lhdr(5)
    allocate(ib_ud(jde1:jde2,ide1:ide2), ib_vd(jde1:jde2,ide1:ide2))
    allocate(ib_cr(jde1:jde2,ide1:ide2), hefc(nspgx,kz), x1(nsteps))
    allocate(bps0(jce1:jce2,ice1:ice2), bps1(jce1:jce2,ice1:ice2))
    allocate(bu0(jde1:jde2,ice1:ice2,1:kz), bu1(jde1:jde2,ice1:ice2,1:kz))
    allocate(bv0(jce1:jce2,ide1:ide2,1:kz), bv1(jce1:jce2,ide1:ide2,1:kz))
    allocate(bt0(jce1:jce2,ice1:ice2,1:kz), bt1(jce1:jce2,ice1:ice2,1:kz))
    allocate(bp0(jce1:jce2,ice1:ice2,1:kz), bp1(jce1:jce2,ice1:ice2,1:kz))
    allocate(bq0(jce1:jce2,ice1:ice2,1:kz), bq1(jce1:jce2,ice1:ice2,1:kz))
    allocate(bc0(jce1:jce2,ice1:ice2,1:kz), bc1(jce1:jce2,ice1:ice2,1:kz))
    allocate(bi0(jce1:jce2,ice1:ice2,1:kz), bi1(jce1:jce2,ice1:ice2,1:kz))
    allocate(tnudge(1:kz))
    read(uin) ib_ud, ib_vd, ib_cr, hefc
    read(uin) bps0, bps1, bu0, bu1, bv0, bv1, bt0, bt1, bp0, bp1, bq0, bq1, bc0, bc1, &
              bi0, bi1, tnudge, x1
  end subroutine read_lbc

  subroutine read_all
    read(uin) fmz, fmzf, rfmzu, rfmzv, bdywtu, bdywtv, bdywtw, zeta, dz
    read(uin) mx, mu, mv, hx, hy, coru, corv, rlat, ffilt
    read(uin) ht, htu, htv
    read(uin) u, v, w, pai, tetav, t, tvirt, qx, ux, vx, p, rho, qsat, ps
  end subroutine read_all

end program moloch_gpu_driver
