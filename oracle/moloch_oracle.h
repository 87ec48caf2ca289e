/*
 * CPU oracle for the MOLOCH dynamical-core step (test infrastructure only).
 *
 * A plain-C, FP64 restatement of the reference's algorithm, following
 * Main/mod_moloch.F90 loop nest by loop nest, with no FMA contraction and
 * Fortran's left-to-right association.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark may call it; the product path never does.
 *
 * Array convention: every 3-D field is stored over the full dot-point extent
 * (jx, iy, nk) in Fortran order (j fastest), 1-based indices mapped to
 * a[((k-1)*iy + (i-1))*jx + (j-1)], with nk = kz or kz+1.  2-D fields are
 * (jx, iy).  Water species qx are nqx consecutive 3-D blocks.  Points a
 * reference array does not own (e.g. cross fields at j = jx) are ignored.
 */
#ifndef MOLOCH_ORACLE_H
#define MOLOCH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
  int jx, iy, kz;        /* dot-point extents and number of zita levels */
  int nqx;               /* number of water species held in qx */
  int ipptls;            /* 0, 1, 2: which species enter Tv and advection */
  int nadv, nsound;      /* mo_nadv, mo_nsound */
  int do_divdamp;        /* mo_divdamp */
  int do_divfilter;      /* mo_divfilter */
  int lrotllr;           /* iproj == 'ROTLLR' */
  int nthreads;          /* OpenMP threads (0 = runtime default) */
  double dx;             /* grid spacing (m) */
  double dtsec;          /* model time step (s) */
  double mo_ztop;        /* model top (m) */
  double mo_a0;          /* gzita shape parameter */
  int ibltyp;            /* 2: TKE carried through the step */
  int ntr;               /* chemical tracers (ichem == 1 when > 0) */
  double tkemin;         /* the PBL scheme's TKE floor */
  int crm;               /* i_crm = 1: doubly periodic (no lateral boundaries) */
  int pad_;
} orc_grid;

typedef struct orc_fields {
  /* prognostic / diagnostic state (mo_atm) */
  double *u, *v, *w, *pai, *tetav, *t, *tvirt, *qx;
  double *ux, *vx, *p, *rho, *qsat, *ps;
  /* tendencies (physics input; zero when physics is off) */
  double *tten, *uten, *vten, *qxten;
  /* static metrics */
  double *fmz, *fmzf, *rfmzu, *rfmzv;       /* kz, kz+1, kz, kz */
  double *bdywtu, *bdywtv, *bdywtw;         /* kz */
  double *zeta, *dz;                        /* kz */
  double *mx, *mu, *mv, *hx, *hy, *coru, *corv;  /* 2-D */
  double *rlat;                             /* 1-D (iy), ROTLLR only */
  double *ffilt;                            /* 1-D (kz) sponge profile */
  /* ibltyp 2: tke, tketen (kz+1); ntr > 0: trac, chiten (ntr blocks of kz) */
  double *tke, *tketen, *trac, *chiten;
} orc_fields;

/* Lateral-boundary inputs of boundary() (Main/mod_moloch.F90:448-529).
 * b0/b1 are the boundary states at the two bracketing times (full extents,
 * as the v3dbound/v2dbound arrays), x1 the time weight (the reference's
 * x1 = (xbctime + dt)*rtb, x0 = 1 - x1).  xfu/xfv/xfc hold hefc(ibnd(j,i),k)
 * of ba_ud / ba_vd / ba_cr where ibnd > 0 and -1 elsewhere. */
typedef struct orc_bdy {
  double x1;
  int present_qc, present_qi;   /* is_present_qc() / is_present_qi() */
  int top_nudge, nztop;         /* mo_top_nudge and its level count */
  double *ps0, *ps1;            /* 2-D */
  double *u0, *u1, *v0, *v1, *t0, *t1, *pai0, *pai1, *qv0, *qv1, *qc0, *qc1, *qi0, *qi1;
  double *xfu, *xfv, *xfc;      /* 3-D relaxation weights */
  double *tnudge;               /* 1-D (kz) top-nudging profile */
  /* spectral nudging (mo_spectral_nudge, Main/mod_bdycod.F90:3844-3960):
   * bvx(jx, 2km), bvy(iy, 2lm) in Fortran order, cnudge(kz); tspectral is
   * the accumulated time of Main/mod_moloch.F90:453 (in/out) */
  int spectral_nudge, km, lm, pad_;
  double dtrad, tspectral;
  double *bvx, *bvy, *cnudge;
} orc_bdy;

/* Inputs and outputs of mkslice's MOLOCH branch (Main/mod_slice.F90:
 * 112-160, 340-382).  ps stands for atms%ps2d; zetaf is the full-level
 * height mo_atm%zetaf (kz+1); xlat the latitude (2-D).  Outputs: pf
 * (kz+1), th3d, rhb3d, wpx3d (kz), rhox2d, tp2d, th700, ptrop, and the
 * levels ktrop, kmxpbl stored as doubles. */
typedef struct orc_slice {
  double rhmin, rhmax, calday, dayspy;
  int icldmstrat, irceideal, ibltyp, pad_;
  const double *zetaf, *xlat;
  double *pf, *th3d, *rhb3d, *wpx3d, *rhox2d, *tp2d, *th700, *ptrop, *ktrop, *kmxpbl;
} orc_slice;

typedef struct orc_mass {
  double drymass, qmass, dryadv, qadv;
} orc_mass;

/* One moloch() call (Main/mod_moloch.F90:312-446) with physics disabled and
 * fixed lateral boundaries: reset_tendencies, dynamical_core, p/rho/qsat,
 * extrapolate_surface_pressure, status_update(dtsec).  Repeated nsteps times.
 * Returns 0, or -1 on allocation failure / bad grid. */
int orc_moloch_steps(const orc_grid *g, orc_fields *f, int nsteps);

/* boundary() only (Main/mod_moloch.F90:448-529): bdyval's MOLOCH branch
 * (Main/mod_bdycod.F90:1638-1866), motopnudge (:4049-4081), morelax
 * (:3962-4031), mospectral_nudge (:3940-3960, every dtrad), uvstagtouvx,
 * temp_to_tvirt and tetav = tvirt/pai.  No tke, no tracers. */
int orc_boundary(const orc_grid *g, orc_fields *f, orc_bdy *b);

/* mkslice, MOLOCH branch (also clamps qx in place, as the reference does
 * through the atms%qxb3d alias). */
int orc_mkslice(const orc_grid *g, orc_fields *f, orc_slice *sl);

/* moloch() with do_apply_bdy: as orc_moloch_steps, with boundary() between
 * dynamical_core and the diagnostics; step n uses x1[n]. */
int orc_moloch_steps_lbc(const orc_grid *g, orc_fields *f, orc_bdy *b,
                         const double *x1, int nsteps);

/* dynamical_core only (Main/mod_moloch.F90:1085-1141). */
int orc_dynamical_core(const orc_grid *g, orc_fields *f);

/* status_update only (Main/mod_moloch.F90:1403-1443). */
int orc_status_update(const orc_grid *g, orc_fields *f, double dtinc);

/* sound() alone (Main/mod_moloch.F90:545-736); stop = 1, 2, 3 return in the
 * first sub-step after zdiv2 is formed, after damping and filter, after the
 * vertical-divergence term.  zdiv2_out (kz) and s_out (kz+1) may be null. */
int orc_sound(const orc_grid *g, orc_fields *f, int stop, double *zdiv2_out,
              double *s_out);

/* Curvature terms of advection (:811-825) on ux, vx. */
int orc_curvature(const orc_grid *g, orc_fields *f, double dta);

/* p, rho, qsat (:348-352) and extrapolate_surface_pressure (:1592-1606). */
int orc_diagnostics(const orc_grid *g, orc_fields *f);

/* init's restart branch for MOLOCH (Main/mod_init.F90:941-955): p, qsat,
 * rho, tvirt = t*(1 + ep1*qv) and tetav = tvirt/pai from the SAV variables. */
int orc_init_restart(const orc_grid *g, orc_fields *f);

/* 0: temp_to_tvirt (:1608-1627), 1: tvirt_to_temp (:1629-1648). */
int orc_tv_convert(const orc_grid *g, orc_fields *f, int which);

/* Plain per-substep pieces, exported so tests can pin them separately. */
int orc_wafone(const orc_grid *g, orc_fields *f, const double *s,
               double *pp, double dta);
int orc_thomas(int n, const double *zu, const double *zd,
               const double *rhs, double wtop, double *wout);

/* Mass budget in the MOLOCH branch of Main/mod_massck.F90:77-160 with the
 * reference's serial k,i,j summation order.  dt is the model time step. */
int orc_massck(const orc_grid *g, const orc_fields *f, orc_mass *m);

/* b(r) of the WAF limiter on local_flow_param(num, den) (:1571-1590, :882). */
double orc_waf_b(double num, double den);

/* xkdamp / xknu profiles (:294-299). */
int orc_damping_profiles(int kz, double *xkdamp, double *xknu);

/* Staggering operators (:1445-1569); see the implementation for `which`. */
int orc_stagger(const orc_grid *g, orc_fields *f, int which);

/* zitaf (kz+1), zitah (kz) and gzita on both (Share/mod_zita.F90:54-119). */
int orc_zita_levels(int kz, double ztop, double a0, double *zitaf,
                    double *zitah, double *gzitaf, double *gzitah);

/* Saturation mixing ratio (Share/pfwsat.inc). */
double orc_pfwsat(double t, double p);

/* Thermodynamic constants as the oracle uses them, by name (NaN if unknown). */
double orc_constant(const char *name);

#ifdef __cplusplus
}
#endif
#endif
