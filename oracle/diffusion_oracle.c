/*
 * CPU oracle of Main/mod_diffusion.F90 (test infrastructure only; see
 * diffusion_oracle.h).  Loop nests follow the reference one by one, with
 * the reference's association and its quirks kept:
 *  - idiffu = 1: the del2 ring is applied side by side (left, right, bottom,
 *    top), so the four ring corners receive it twice (:301-367, :537-573);
 *  - diffu_x3df, idiffu = 1: the left / right ring runs over k = 1..kz and
 *    the bottom / top ring and the interior over 1..kz+1 (:532-572);
 *  - idiffu = 3: only the column j = jci2 (jdi2 for diffu_d) is diffused
 *    (:411, :593, :719, :868), and the flux differences use f while the
 *    limiter tests use f / msfd (diffu_x);
 *  - diffu_x3df, idiffu = 3, level kz+1: the reference reads
 *    xkc(:,:,kz+1), one level past xkc's allocation (1:kz).  Under
 *    idiffu = 3 every level of xkc and xkcf holds diff_6th_coef * psb, so
 *    this restatement reads xkcf(:,:,kz+1), which holds that value.
 */
#include "diffusion_oracle.h"

#include <math.h>
#include <stddef.h>

#define D3(a, j, i, k) \
  (a)[((size_t)((k)-1) * (size_t)d->iy + (size_t)((i)-1)) * (size_t)d->jx + (size_t)((j)-1)]
#define D2(a, j, i) (a)[(size_t)((i)-1) * (size_t)d->jx + (size_t)((j)-1)]

/* Share/mod_constants.F90 */
static const double c_egrav = 9.80665;
static const double c_vonkar = 0.4;

/* 9-point Laplacian weights (LeVeque eq. 3.17), del4 weights, 6th-order
 * flux weights: Main/mod_diffusion.F90:65-83 */
static const double o4_c1 = 4.0 / 6.0, o4_c2 = 1.0 / 6.0, o4_c3 = -20.0 / 6.0;
static const double z4_c1 = 1.0, z4_c2 = -4.0, z4_c3 = 12.0;
static const double h4_c1 = 10.0, h4_c2 = -5.0, h4_c3 = 1.0;
static const double diff_6th_factor = 0.12;

static int bad(const orc_diff *d) {
  return !d || d->jx < 8 || d->iy < 8 || d->kz < 1 || d->idiffu < 1 || d->idiffu > 3 ||
         d->idynamic < 1 || d->idynamic > 2 || !(d->dx > 0.0) || !(d->dtsec > 0.0);
}

/* initialize_diffusion :110-125, :152-153 */
int orc_diff_scalars(const orc_diff *d, double out[4]) {
  if (bad(d)) return -1;
  const double dxsq = d->dx * d->dx;
  double xkhmax = dxsq / (64.0 * d->dtsec);
  const double dydc = d->adyndif * c_vonkar * c_vonkar * d->dx * 0.25;
  double xkhz;
  if (d->idynamic == 1) {
    xkhz = d->ckh * 1.5e-3 * dxsq / d->dtsec;
  } else {
    xkhz = d->ckh * d->dx;
    xkhmax = 2.0 * xkhmax;
  }
  out[0] = xkhz;
  out[1] = xkhmax;
  out[2] = dydc;
  out[3] = diff_6th_factor * 0.015625 / (2.0 * d->dtsec);
  return 0;
}

/* initialize_diffusion :131-146 (hgfact on jce x ice, corrected on jci x ici) */
int orc_diff_init(orc_diff *d) {
  double s[4];
  if (orc_diff_scalars(d, s)) return -1;
  if (d->idiffu >= 3) return 0;
  const int jx = d->jx, iy = d->iy;
  const double xkhz = s[0], regrav = 1.0 / c_egrav;
  for (int i = 1; i <= iy - 1; ++i)
    for (int j = 1; j <= jx - 1; ++j) D2(d->hgfact, j, i) = xkhz;
  if (d->diffu_hgtf == 1) {
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j) {
        const double h = D2(d->ht, j, i);
        const double hg1 = fabs((h - D2(d->ht, j, i - 1)) / d->dx);
        const double hg2 = fabs((h - D2(d->ht, j, i + 1)) / d->dx);
        const double hg3 = fabs((h - D2(d->ht, j - 1, i)) / d->dx);
        const double hg4 = fabs((h - D2(d->ht, j + 1, i)) / d->dx);
        double m = hg1;
        if (hg2 > m) m = hg2;
        if (hg3 > m) m = hg3;
        if (hg4 > m) m = hg4;
        const double hgmax = m * regrav * 1.0e3;
        D2(d->hgfact, j, i) = xkhz / (1.0 + hgmax * hgmax);
      }
  }
  return 0;
}

/* calc_coeff :166-248 */
int orc_calc_coeff(orc_diff *d) {
  double s[4];
  if (orc_diff_scalars(d, s)) return -1;
  const int jx = d->jx, iy = d->iy, kz = d->kz;
  const double xkhmax = s[1], dydc = s[2], coef6 = s[3];
  if (d->idiffu == 3) {
    for (int k = 1; k <= kz; ++k)
      for (int i = 2; i <= iy - 2; ++i)
        for (int j = 2; j <= jx - 2; ++j) D3(d->xkc, j, i, k) = coef6 * D2(d->pc, j, i);
    for (int k = 1; k <= kz + 1; ++k)
      for (int i = 2; i <= iy - 2; ++i)
        for (int j = 2; j <= jx - 2; ++j) D3(d->xkcf, j, i, k) = coef6 * D2(d->pc, j, i);
    for (int k = 1; k <= kz; ++k)
      for (int i = 2; i <= iy - 1; ++i)
        for (int j = 2; j <= jx - 1; ++j) D3(d->xkd, j, i, k) = coef6 * D2(d->pd, j, i);
    return 0;
  }
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j) D3(d->xkc, j, i, k) = 0.0;
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 1; ++i)
      for (int j = 2; j <= jx - 1; ++j) D3(d->xkd, j, i, k) = 0.0;
  for (int k = 1; k <= kz + 1; ++k)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j) D3(d->xkcf, j, i, k) = 0.0;
  const double *ud = d->ud, *vd = d->vd;
  for (int k = 1; k <= kz; ++k)
    for (int i = 1; i <= iy - 1; ++i)
      for (int j = 1; j <= jx - 1; ++j) {
        const double dudx = D3(ud, j + 1, i, k) + D3(ud, j + 1, i + 1, k) -
                            D3(ud, j, i, k) - D3(ud, j, i + 1, k);
        const double dvdx = D3(vd, j + 1, i, k) + D3(vd, j + 1, i + 1, k) -
                            D3(vd, j, i, k) - D3(vd, j, i + 1, k);
        const double dudy = D3(ud, j, i + 1, k) + D3(ud, j + 1, i + 1, k) -
                            D3(ud, j, i, k) - D3(ud, j + 1, i, k);
        const double dvdy = D3(vd, j, i + 1, k) + D3(vd, j + 1, i + 1, k) -
                            D3(vd, j, i, k) - D3(vd, j + 1, i, k);
        double duv;
        if (d->idynamic == 1) {
          duv = sqrt((dudx - dvdy) * (dudx - dvdy) + (dvdx + dudy) * (dvdx + dudy));
        } else {
          const double dwdz = D3(d->wx, j, i, k) - D3(d->wx, j, i, k + 1);
          double a = (dudx - dvdy) * (dudx - dvdy) + (dvdx + dudy) * (dvdx + dudy) -
                     dwdz * dwdz;
          if (!(a > 0.0)) a = 0.0;
          duv = sqrt(a);
        }
        const double x = D2(d->hgfact, j, i) + dydc * duv;
        D3(d->xkc, j, i, k) = x < xkhmax ? x : xkhmax;
      }
  for (int i = 2; i <= iy - 2; ++i)
    for (int j = 2; j <= jx - 2; ++j) D3(d->xkcf, j, i, 1) = D3(d->xkc, j, i, 1);
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j) D3(d->xkcf, j, i, k + 1) = D3(d->xkc, j, i, k);
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 1; ++i)
      for (int j = 2; j <= jx - 1; ++j)
        D3(d->xkd, j, i, k) = 0.25 * (D3(d->xkc, j, i, k) + D3(d->xkc, j - 1, i - 1, k) +
                                      D3(d->xkc, j - 1, i, k) + D3(d->xkc, j, i - 1, k));
  const double rdxsq = 1.0 / (d->dx * d->dx);
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j)
        D3(d->xkc, j, i, k) = D3(d->xkc, j, i, k) * rdxsq * D2(d->pc, j, i);
  for (int k = 1; k <= kz + 1; ++k)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j)
        D3(d->xkcf, j, i, k) = D3(d->xkcf, j, i, k) * rdxsq * D2(d->pc, j, i);
  for (int k = 1; k <= kz; ++k)
    for (int i = 2; i <= iy - 1; ++i)
      for (int j = 2; j <= jx - 1; ++j)
        D3(d->xkd, j, i, k) = D3(d->xkd, j, i, k) * rdxsq * D2(d->pd, j, i);
  return 0;
}

/* ---- diffu_d :263-511 ---------------------------------------------------- */

/* f / msfd at a dot point */
#define UM(f, j, i, k) (D3(f, j, i, k) / D2(d->msfd, j, i))

/* 13-point del4 operand of the interior (:283-293) */
static double del4_d(const orc_diff *d, const double *u, int j, int i, int k) {
  return z4_c1 * (UM(u, j + 2, i, k) + UM(u, j - 2, i, k) + UM(u, j, i + 2, k) +
                  UM(u, j, i - 2, k)) +
         z4_c2 * (UM(u, j + 1, i, k) + UM(u, j - 1, i, k) + UM(u, j, i + 1, k) +
                  UM(u, j, i - 1, k)) +
         z4_c3 * (UM(u, j, i, k));
}

/* 5-point del2 operand of the ring (:306-311) */
static double del2_d(const orc_diff *d, const double *u, int j, int i, int k) {
  return z4_c1 * (UM(u, j + 1, i, k) + UM(u, j - 1, i, k) + UM(u, j, i + 1, k) +
                  UM(u, j, i - 1, k)) +
         z4_c2 * (UM(u, j, i, k));
}

/* 9-point operand (:374-383) */
static double nine_d(const orc_diff *d, const double *u, int j, int i, int k) {
  return o4_c1 * (UM(u, j + 1, i, k) + UM(u, j - 1, i, k) + UM(u, j, i + 1, k) +
                  UM(u, j, i - 1, k)) +
         o4_c2 * (UM(u, j + 1, i + 1, k) + UM(u, j - 1, i - 1, k) + UM(u, j - 1, i + 1, k) +
                  UM(u, j + 1, i - 1, k)) +
         o4_c3 * (UM(u, j, i, k));
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* flux-limited 6th-order operand of diffu_d (:399-460): f / msfd everywhere */
static double flux6_d(const orc_diff *d, const double *u, int j, int i, int k) {
  const int im1 = imax(i - 1, 1), im2 = imax(i - 2, 1), im3 = imax(i - 3, 1);
  const int ip1 = imin(i + 1, d->iy), ip2 = imin(i + 2, d->iy), ip3 = imin(i + 3, d->iy);
  const int jm1 = imax(j - 1, 1), jm2 = imax(j - 2, 1), jm3 = imax(j - 3, 1);
  const int jp1 = imin(j + 1, d->jx), jp2 = imin(j + 2, d->jx), jp3 = imin(j + 3, d->jx);
  double x0 = h4_c1 * (UM(u, j, i, k) - UM(u, jm1, i, k)) +
              h4_c2 * (UM(u, jp1, i, k) - UM(u, jm2, i, k)) +
              h4_c3 * (UM(u, jp2, i, k) - UM(u, jm3, i, k));
  if (x0 * (UM(u, j, i, k) - UM(u, jm1, i, k)) <= 0.0) x0 = 0.0;
  double x1 = h4_c1 * (UM(u, jp1, i, k) - UM(u, j, i, k)) +
              h4_c2 * (UM(u, jp2, i, k) - UM(u, jm1, i, k)) +
              h4_c3 * (UM(u, jp3, i, k) - UM(u, jm2, i, k));
  if (x1 * (UM(u, jp1, i, k) - UM(u, j, i, k)) <= 0.0) x1 = 0.0;
  double y0 = h4_c1 * (UM(u, j, i, k) - UM(u, j, im1, k)) +
              h4_c2 * (UM(u, j, ip1, k) - UM(u, j, im2, k)) +
              h4_c3 * (UM(u, j, ip2, k) - UM(u, j, im3, k));
  if (y0 * (UM(u, j, i, k) - UM(u, j, im1, k)) <= 0.0) y0 = 0.0;
  double y1 = h4_c1 * (UM(u, j, ip1, k) - UM(u, j, i, k)) +
              h4_c2 * (UM(u, j, ip2, k) - UM(u, j, im1, k)) +
              h4_c3 * (UM(u, j, ip3, k) - UM(u, j, im2, k));
  if (y1 * (UM(u, j, ip1, k) - UM(u, j, i, k)) <= 0.0) y1 = 0.0;
  return (x1 - x0) + (y1 - y0);
}

int orc_diffu_d(const orc_diff *d, double *uten, double *vten, const double *u,
                const double *v) {
  if (bad(d)) return -1;
  const int jx = d->jx, iy = d->iy, kz = d->kz;
  const int jdi1 = 2, jdi2 = jx - 1, idi1 = 2, idi2 = iy - 1;
  if (d->idiffu == 1) {
    for (int k = 1; k <= kz; ++k)
      for (int i = 3; i <= iy - 2; ++i)
        for (int j = 3; j <= jx - 2; ++j) {
          D3(uten, j, i, k) = D3(uten, j, i, k) - D3(d->xkd, j, i, k) * del4_d(d, u, j, i, k);
          D3(vten, j, i, k) = D3(vten, j, i, k) - D3(d->xkd, j, i, k) * del4_d(d, v, j, i, k);
        }
    const int side_j[2] = {jdi1, jdi2}, side_i[2] = {idi1, idi2};
    for (int s = 0; s < 2; ++s) {          /* left, right */
      const int j = side_j[s];
      for (int k = 1; k <= kz; ++k)
        for (int i = idi1; i <= idi2; ++i) {
          D3(uten, j, i, k) = D3(uten, j, i, k) + D3(d->xkd, j, i, k) * del2_d(d, u, j, i, k);
          D3(vten, j, i, k) = D3(vten, j, i, k) + D3(d->xkd, j, i, k) * del2_d(d, v, j, i, k);
        }
    }
    for (int s = 0; s < 2; ++s) {          /* bottom, top */
      const int i = side_i[s];
      for (int k = 1; k <= kz; ++k)
        for (int j = jdi1; j <= jdi2; ++j) {
          D3(uten, j, i, k) = D3(uten, j, i, k) + D3(d->xkd, j, i, k) * del2_d(d, u, j, i, k);
          D3(vten, j, i, k) = D3(vten, j, i, k) + D3(d->xkd, j, i, k) * del2_d(d, v, j, i, k);
        }
    }
  } else if (d->idiffu == 2) {
    for (int k = 1; k <= kz; ++k)
      for (int i = idi1; i <= idi2; ++i)
        for (int j = jdi1; j <= jdi2; ++j) {
          D3(uten, j, i, k) = D3(uten, j, i, k) + D3(d->xkd, j, i, k) * nine_d(d, u, j, i, k);
          D3(vten, j, i, k) = D3(vten, j, i, k) + D3(d->xkd, j, i, k) * nine_d(d, v, j, i, k);
        }
  } else {
    for (int k = 1; k <= kz; ++k)
      for (int i = idi1; i <= idi2; ++i) {
        const int j = jdi2;
        D3(uten, j, i, k) = D3(uten, j, i, k) + D3(d->xkd, j, i, k) * flux6_d(d, u, j, i, k);
        D3(vten, j, i, k) = D3(vten, j, i, k) + D3(d->xkd, j, i, k) * flux6_d(d, v, j, i, k);
      }
  }
  return 0;
}

/* ---- diffu_x :513-913 ---------------------------------------------------- */

#define F(j, i, k) D3(f, j, i, k)

static double del4_x(const orc_diff *d, const double *f, int j, int i, int k) {
  return z4_c1 * (F(j + 2, i, k) + F(j - 2, i, k) + F(j, i + 2, k) + F(j, i - 2, k)) +
         z4_c2 * (F(j + 1, i, k) + F(j - 1, i, k) + F(j, i + 1, k) + F(j, i - 1, k)) +
         z4_c3 * (F(j, i, k));
}

static double del2_x(const orc_diff *d, const double *f, int j, int i, int k) {
  return z4_c1 * (F(j + 1, i, k) + F(j - 1, i, k) + F(j, i + 1, k) + F(j, i - 1, k)) +
         z4_c2 * (F(j, i, k));
}

static double nine_x(const orc_diff *d, const double *f, int j, int i, int k) {
  return o4_c1 * (F(j + 1, i, k) + F(j - 1, i, k) + F(j, i + 1, k) + F(j, i - 1, k)) +
         o4_c2 * (F(j + 1, i + 1, k) + F(j - 1, i - 1, k) + F(j - 1, i + 1, k) +
                  F(j + 1, i - 1, k)) +
         o4_c3 * (F(j, i, k));
}

/* flux-limited 6th-order operand of diffu_x (:581-630): differences of f,
 * limiter tests on f / msfd; indices clamped to the cross extent */
static double flux6_x(const orc_diff *d, const double *f, int j, int i, int k) {
  const int iym1 = d->iy - 1, jxm1 = d->jx - 1;
  const int im1 = imax(i - 1, 1), im2 = imax(i - 2, 1), im3 = imax(i - 3, 1);
  const int ip1 = imin(i + 1, iym1), ip2 = imin(i + 2, iym1), ip3 = imin(i + 3, iym1);
  const int jm1 = imax(j - 1, 1), jm2 = imax(j - 2, 1), jm3 = imax(j - 3, 1);
  const int jp1 = imin(j + 1, jxm1), jp2 = imin(j + 2, jxm1), jp3 = imin(j + 3, jxm1);
  double x0 = h4_c1 * (F(j, i, k) - F(jm1, i, k)) + h4_c2 * (F(jp1, i, k) - F(jm2, i, k)) +
              h4_c3 * (F(jp2, i, k) - F(jm3, i, k));
  if (x0 * (UM(f, j, i, k) - UM(f, jm1, i, k)) <= 0.0) x0 = 0.0;
  double x1 = h4_c1 * (F(jp1, i, k) - F(j, i, k)) + h4_c2 * (F(jp2, i, k) - F(jm1, i, k)) +
              h4_c3 * (F(jp3, i, k) - F(jm2, i, k));
  if (x1 * (UM(f, jp1, i, k) - UM(f, j, i, k)) <= 0.0) x1 = 0.0;
  double y0 = h4_c1 * (F(j, i, k) - F(j, im1, k)) + h4_c2 * (F(j, ip1, k) - F(j, im2, k)) +
              h4_c3 * (F(j, ip2, k) - F(j, im3, k));
  if (y0 * (UM(f, j, i, k) - UM(f, j, im1, k)) <= 0.0) y0 = 0.0;
  double y1 = h4_c1 * (F(j, ip1, k) - F(j, i, k)) + h4_c2 * (F(j, ip2, k) - F(j, im1, k)) +
              h4_c3 * (F(j, ip3, k) - F(j, im2, k));
  if (y1 * (UM(f, j, ip1, k) - UM(f, j, i, k)) <= 0.0) y1 = 0.0;
  return (x1 - x0) + (y1 - y0);
}

/* one 3-D field: form 0 (x3d), 1 (x3df), 2 (x4d3d) */
static void diffu_x_one(const orc_diff *d, int form, double *ften, const double *f,
                        double fac) {
  const int jx = d->jx, iy = d->iy, kz = d->kz;
  const int jci1 = 2, jci2 = jx - 2, ici1 = 2, ici2 = iy - 2;
  const int nk = form == 1 ? kz + 1 : kz;
  const double *xk = form == 1 ? d->xkcf : d->xkc;
  /* the coefficient factor of each form: x3d none, x3df and x4d3d fac */
#define KF(j, i, k) (form == 0 ? D3(xk, j, i, k) : fac * D3(xk, j, i, k))
  if (d->idiffu == 1) {
    for (int k = 1; k <= nk; ++k)
      for (int i = 3; i <= iy - 3; ++i)
        for (int j = 3; j <= jx - 3; ++j)
          D3(ften, j, i, k) = D3(ften, j, i, k) - KF(j, i, k) * del4_x(d, f, j, i, k);
    const int nk_lr = kz;                 /* left / right rings: 1..kz in every form */
    const int side_j[2] = {jci1, jci2}, side_i[2] = {ici1, ici2};
    for (int s = 0; s < 2; ++s) {
      const int j = side_j[s];
      for (int k = 1; k <= nk_lr; ++k)
        for (int i = ici1; i <= ici2; ++i)
          D3(ften, j, i, k) = D3(ften, j, i, k) + KF(j, i, k) * del2_x(d, f, j, i, k);
    }
    for (int s = 0; s < 2; ++s) {
      const int i = side_i[s];
      for (int k = 1; k <= nk; ++k)
        for (int j = jci1; j <= jci2; ++j)
          D3(ften, j, i, k) = D3(ften, j, i, k) + KF(j, i, k) * del2_x(d, f, j, i, k);
    }
  } else if (d->idiffu == 2) {
    for (int k = 1; k <= nk; ++k)
      for (int i = ici1; i <= ici2; ++i)
        for (int j = jci1; j <= jci2; ++j)
          D3(ften, j, i, k) = D3(ften, j, i, k) + KF(j, i, k) * nine_x(d, f, j, i, k);
  } else {
    for (int k = 1; k <= nk; ++k)
      for (int i = ici1; i <= ici2; ++i) {
        const int j = jci2;
        /* x3d: xkc; x3df: fac * xkc (level kz+1 from xkcf); x4d3d: xkc, no fac */
        const double xkv = (form == 1 && k == kz + 1) ? D3(d->xkcf, j, i, k)
                                                       : D3(d->xkc, j, i, k);
        const double kf = form == 1 ? fac * xkv : xkv;
        D3(ften, j, i, k) = D3(ften, j, i, k) + kf * flux6_x(d, f, j, i, k);
      }
  }
#undef KF
}

int orc_diffu_x(const orc_diff *d, int form, double *ften, const double *f, int n1,
                int n2, double fac) {
  if (bad(d) || form < 0 || form > 2) return -1;
  if (form != 2) {
    diffu_x_one(d, form, ften, f, fac);
    return 0;
  }
  if (n1 < 1 || n2 < n1) return -1;
  const size_t blk = (size_t)d->jx * (size_t)d->iy * (size_t)d->kz;
  for (int n = n1; n <= n2; ++n)
    diffu_x_one(d, 2, ften + (size_t)(n - 1) * blk, f + (size_t)(n - 1) * blk, fac);
  return 0;
}
