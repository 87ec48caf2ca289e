/*
 * CPU oracle for the horizontal-diffusion operator of the MM cores
 * (Main/mod_diffusion.F90), test infrastructure only: initialize_diffusion
 * (:99-163), calc_coeff (:166-248), diffu_d (:263-511) and the diffu_x
 * family (:513-913), restated in plain C, FP64, without FMA contraction and
 * with Fortran's left-to-right association, for one domain that owns every
 * lateral boundary (ma%has_bdy* all true, so the *ga extents equal the
 * plain ones and the halo exchange of xkc is a no-op).
 *
 * Array convention as in moloch_oracle.h: 3-D arrays over (jx, iy, nk),
 * j fastest, a[((k-1)*iy + (i-1))*jx + (j-1)]; 2-D arrays over (jx, iy);
 * multi-species arrays are consecutive 3-D blocks.
 */
#ifndef DIFFUSION_ORACLE_H
#define DIFFUSION_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_diff {
  int jx, iy, kz;
  int idiffu;           /* 1 del4 + del2 ring, 2 nine-point, 3 flux-limited 6th order */
  int idynamic;         /* 1 hydrostatic, 2 non-hydrostatic (coefficient uses wx) */
  int diffu_hgtf;       /* topographic reduction of the background coefficient */
  int pad_[2];
  double dx, dtsec, ckh, adyndif;
  const double *ht;     /* mddom%ht (jx, iy) */
  const double *msfd;   /* mddom%msfd (jx, iy), dot points */
  const double *pc;     /* sfs%psb (jx, iy) */
  const double *pd;     /* sfs%psdotb (jx, iy) */
  const double *ud, *vd;   /* atms%ubd3d, vbd3d (kz) */
  const double *wx;        /* atms%wb3d (kz+1) */
  double *hgfact;          /* (jx, iy) */
  double *xkc, *xkd;       /* (kz) */
  double *xkcf;            /* (kz+1) */
} orc_diff;

/* initialize_diffusion: hgfact (idiffu < 3). */
int orc_diff_init(orc_diff *d);
/* calc_coeff: xkc, xkd, xkcf. */
int orc_calc_coeff(orc_diff *d);
/* diffu_d(uten, vten, u, v), kz levels, dot points. */
int orc_diffu_d(const orc_diff *d, double *uten, double *vten, const double *u,
                const double *v);
/* diffu_x: form 0 = diffu_x3d(ften, f) over kz levels; form 1 =
 * diffu_x3df(ften, f, fac) over kz+1 levels; form 2 = diffu_x4d(ften, f,
 * n1, n2, fac) over kz levels and species n1..n2 (1-based; ften and f hold
 * n2 species). */
int orc_diffu_x(const orc_diff *d, int form, double *ften, const double *f, int n1,
                int n2, double fac);
/* The scalars initialize_diffusion derives: xkhz, xkhmax, dydc, diff_6th_coef. */
int orc_diff_scalars(const orc_diff *d, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
