! Probe driver for the reference's leaf modules, compiled verbatim from
! Share/mod_realkinds.F90, Share/mod_intkinds.F90, Share/mod_constants.F90,
! Share/mod_zita.F90 and Share/pfwsat.inc (included the same way
! Main/mod_moloch.F90:157 includes it).  This file is test infrastructure:
! it prints the bit patterns of the physical constants, of the zita
! closed forms and of pfwsat so that the oracle's restatement of them can
! be pinned against the reference's own compiled code.
!
! Output: one record per line, "<tag> <index...> <hex64>".
module ref_probe_sat
  use mod_realkinds
  use mod_constants
  implicit none
  private
  public :: pfwsat
  contains
#include <pfwsat.inc>
end module ref_probe_sat

program ref_probe
  use, intrinsic :: iso_fortran_env, only : int64
  use mod_realkinds
  use mod_constants
  use mod_zita
  use ref_probe_sat
  implicit none
  integer, parameter :: nkz = 3
  integer, dimension(nkz) :: kzlist = [ 10, 23, 41 ]
  real(rk8), dimension(3) :: ztops = [ 30000.0_rk8, 28000.0_rk8, 25000.0_rk8 ]
  real(rk8), dimension(3) :: zhs = [ 8000.0_rk8, 7500.0_rk8, 6000.0_rk8 ]
  real(rk8), dimension(3) :: a0s = [ 0.0_rk8, 0.0_rk8, 0.5_rk8 ]
  real(rk8), dimension(3) :: geos = [ 0.0_rk8, 4903.325_rk8, 19613.3_rk8 ]
  real(rk8), dimension(12) :: temps = [ 180.0_rk8, 198.15_rk8, 210.5_rk8, &
       230.0_rk8, 250.25_rk8, 263.0_rk8, 273.15_rk8, 273.16_rk8, &
       285.0_rk8, 300.0_rk8, 320.0_rk8, 380.0_rk8 ]
  real(rk8), dimension(6) :: press = [ 500.0_rk8, 2000.0_rk8, 20000.0_rk8, &
       50000.0_rk8, 85000.0_rk8, 101325.0_rk8 ]
  real(rk8), allocatable, dimension(:) :: zf, zh, gf, gh
  integer :: n, k, kz, m, ig, it, ip
  real(rk8) :: x

  ! Constants used on the MOLOCH step (Share/mod_constants.F90)
  call pr1('egrav', egrav)
  call pr1('rgas', rgas)
  call pr1('cpd', cpd)
  call pr1('cvd', cvd)
  call pr1('rdrcv', rdrcv)
  call pr1('cpovr', cpovr)
  call pr1('rovcp', rovcp)
  call pr1('govr', govr)
  call pr1('govcp', govcp)
  call pr1('regrav', regrav)
  call pr1('ep1', ep1)
  call pr1('ep2', ep2)
  call pr1('tzero', tzero)
  call pr1('p00', p00)
  call pr1('lrate', lrate)
  call pr1('eomeg2', eomeg2)
  call pr1('degrad', degrad)
  call pr1('mathpi', mathpi)
  call pr1('rwat', rwat)
  call pr1('earthrad', earthrad)

  ! Zita vertical coordinate (Share/mod_zita.F90)
  do n = 1, nkz
    kz = kzlist(n)
    allocate(zf(kz+1), zh(kz), gf(kz+1), gh(kz))
    do m = 1, size(ztops)
      call model_zitaf(zf,ztops(m))
      call model_zitah(zh,ztops(m))
      gf = gzita(zf,ztops(m),a0s(m))
      gh = gzita(zh,ztops(m),a0s(m))
      do k = 1, kz+1
        call pr3('zitaf', kz, m, k, zf(k))
        call pr3('gzitaf', kz, m, k, gf(k))
        call pr3('bzitaf', kz, m, k, bzita(zf(k),ztops(m),zhs(m)))
      end do
      do k = 1, kz
        call pr3('zitah', kz, m, k, zh(k))
        call pr3('gzitah', kz, m, k, gh(k))
      end do
      do ig = 1, size(geos)
        do k = 1, kz+1
          x = md_fmz(zf(k),geos(ig),ztops(m),zhs(m),a0s(m))
          call pr4('fmzf', kz, m, ig, k, x)
          x = md_zeta(zf(k),geos(ig),ztops(m),zhs(m),a0s(m))
          call pr4('zetaf', kz, m, ig, k, x)
        end do
        do k = 1, kz
          x = md_fmz(zh(k),geos(ig),ztops(m),zhs(m),a0s(m))
          call pr4('fmzh', kz, m, ig, k, x)
          x = md_zeta(zh(k),geos(ig),ztops(m),zhs(m),a0s(m))
          call pr4('zetah', kz, m, ig, k, x)
        end do
      end do
    end do
    deallocate(zf, zh, gf, gh)
  end do

  ! Saturation mixing ratio (Share/pfwsat.inc)
  do it = 1, size(temps)
    do ip = 1, size(press)
      call pr3('pfwsat', it, ip, 0, pfwsat(temps(it),press(ip)))
    end do
  end do

  contains

  subroutine pr1(tag, v)
    character(len=*), intent(in) :: tag
    real(rk8), intent(in) :: v
    write(*,'(a,1x,z16.16)') tag, transfer(v, 0_int64)
  end subroutine pr1

  subroutine pr3(tag, a, b, c, v)
    character(len=*), intent(in) :: tag
    integer, intent(in) :: a, b, c
    real(rk8), intent(in) :: v
    write(*,'(a,3(1x,i0),1x,z16.16)') tag, a, b, c, transfer(v, 0_int64)
  end subroutine pr3

  subroutine pr4(tag, a, b, c, d, v)
    character(len=*), intent(in) :: tag
    integer, intent(in) :: a, b, c, d
    real(rk8), intent(in) :: v
    write(*,'(a,4(1x,i0),1x,z16.16)') tag, a, b, c, d, transfer(v, 0_int64)
  end subroutine pr4

end program ref_probe
