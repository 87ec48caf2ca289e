/*
 * CPU oracle for the MOLOCH dynamical-core step.  TEST INFRASTRUCTURE ONLY:
 * tests, the smoke check and the CPU-baseline leg of the benchmark call it;
 * the product (regcm_amd) never links or calls it.
 *
 * Plain-C FP64 restatement of Main/mod_moloch.F90 (ICTP RegCM 5.0.0).  Each
 * routine cites the reference lines it follows.  Arithmetic keeps Fortran's
 * left-to-right association and explicit parentheses; the file must be
 * compiled with -ffp-contract=off so that no multiply-add is fused.
 *
 * Single-domain, non-band extents (Main/mod_atm_interface.F90:231-287 with
 * every edge a physical boundary):
 *   jde = 1..jx     jdi = 2..jx-1   jdii = 3..jx-2
 *   jce = 1..jx-1   jci = 2..jx-2
 *   ide = 1..iy     idi = 2..iy-1   idii = 3..iy-2
 *   ice = 1..iy-1   ici = 2..iy-2
 * All ga/gb halos are empty, so the mpplib exchanges are no-ops
 * (Main/mpplib/mod_mppparam.F90:1262-1316) and are omitted.
 *
 * Doubly periodic domain (i_crm = 1, which implies i_band): one process is
 * its own neighbour on every side (mod_mppparam.F90:1298-1309), no edge is a
 * physical boundary, and the cross and dot extents are both 1..jx, 1..iy
 * (:1290-1297, mod_atm_interface.F90:231-287):
 *   jde = jdi = jdii = jce = jci = 1..jx,  ide = ... = ici = 1..iy
 * Each exchange copies the current values of the opposite side into the
 * halos, and every halo the step reads is exchanged after the last write of
 * its field and before the read (u, v at :570-571 and :1532-1533, tetav :562,
 * zdiv2 :535 and :745, pai :673, wz :924, p0 :955/:1012, ux, vx :1485-1486).
 * So a halo read is the periodic partner's current value, and the oracle
 * reads it there directly: A3/A2 wrap their indices, and the edge formulas
 * guarded by has_bdy* are skipped.  The upwind clamps widen by two
 * (init_moloch :284-293).
 */
#include "moloch_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: Share/mod_constants.F90 (non-RCEMIP branch) ------------ */
static const double c_egrav = 9.80665;                     /* :226-230 */
static const double c_navgdr = 6.02214076e23;
static const double c_boltzk = 1.3806490e-23;
static const double c_amd = 28.96454;
static const double c_amw = 18.01528;
#define c_rgasmol (c_navgdr * c_boltzk)
#define c_rgas ((c_rgasmol / c_amd) * 1000.0)
#define c_rwat ((c_rgasmol / c_amw) * 1000.0)
#define c_cpd (3.5 * c_rgas)
#define c_cvd (2.5 * c_rgas)
#define c_regrav (1.0 / c_egrav)
#define c_rcpd (1.0 / c_cpd)
#define c_rovcp (c_rgas * c_rcpd)
#define c_rdrcv (c_rgas / c_cvd)
#define c_cpovr (c_cpd / c_rgas)
#define c_govr (c_egrav / c_rgas)
#define c_govcp (c_egrav / c_cpd)
#define c_ep1 (c_amd / c_amw - 1.0)
#define c_ep2 (c_amw / c_amd)
static const double c_tzero = 273.15;
static const double c_p00 = 1.0e5;
static const double c_lrate = 0.00649;
static const double c_mathpi = 3.14159265358979323846;
#define c_degrad (c_mathpi / 180.0)
static const double c_earthrad = 6.371229e6;
#define c_rearthrad (1.0 / c_earthrad)
static const double c_eomeg = 7.2921159e-5;
#define c_eomeg2 (2.0 * c_eomeg)

/* qxcheckval / qxzeroval: Main/mpplib/mod_runparams.F90:184-192 */
static const double qxcheckval[10] = {1.0e-8, 1.0e-16, 1.0e-16, 1.0e-16,
                                      1.0e-16, 1.0e-16, 1.0e-16, 1.0e10,
                                      100.0, 0.01};
static const double qxzeroval[10] = {1.0e-8, 0.0, 0.0, 0.0, 0.0,
                                     0.0, 0.0, 1.0e10, 100.0, 0.01};

double orc_constant(const char *name) {
  struct { const char *n; double v; } tab[] = {
      {"egrav", c_egrav},   {"rgas", c_rgas},     {"cpd", c_cpd},
      {"cvd", c_cvd},       {"rdrcv", c_rdrcv},   {"cpovr", c_cpovr},
      {"rovcp", c_rovcp},   {"govr", c_govr},     {"govcp", c_govcp},
      {"regrav", c_regrav}, {"ep1", c_ep1},       {"ep2", c_ep2},
      {"tzero", c_tzero},   {"p00", c_p00},       {"lrate", c_lrate},
      {"eomeg2", c_eomeg2}, {"degrad", c_degrad}, {"mathpi", c_mathpi},
      {"rwat", c_rwat},     {"earthrad", c_earthrad},
  };
  for (size_t n = 0; n < sizeof(tab) / sizeof(tab[0]); ++n)
    if (strcmp(tab[n].n, name) == 0) return tab[n].v;
  return NAN;
}

static inline double fmaxd(double a, double b) { return a > b ? a : b; }
static inline double fmind(double a, double b) { return a < b ? a : b; }

/* ---- context: module variables of mod_moloch plus scratch --------------- */
typedef struct ctx {
  const orc_grid *g;
  orc_fields *f;
  int jx, iy, kz, kzp1, kzm1;
  int jde1, jde2, jdi1, jdi2, jdii1, jdii2;
  int jce1, jce2, jci1, jci2;
  int ide1, ide2, idi1, idi2, idii1, idii2;
  int ice1, ice2, ici1, ici2;
  int jmin, jmax, imin, imax;
  int crm;                       /* doubly periodic: indices wrap */
  double rdx, rdzita, dtstepa, dtsound;
  /* 1-D */
  double *gzitak, *gzitakh, *xkdamp, *xknu, *zita, *zitah;
  /* 2-D derived */
  double *mx2, *rmx, *rmu, *rmv;
  /* 3-D scratch (all zero-filled, as getmem does: Share/mod_space.F90:309) */
  double *s, *wwkw, *tetavf, *zdiv2, *wz, *p0, *wfw, *wx, *zpby, *zpbw;
  double *ud, *vd, *laplacian;
  double *tkex;                  /* ibltyp 2: TKE on half levels */
  int probe_stop;                /* orc_sound: phase after which sound returns */
} ctx;

#define P2(c) ((size_t)(c)->jx * (size_t)(c)->iy)
/* periodic partner of an index within two points of the domain */
static inline int wrap_index(int crm, int j, int n) {
  return crm ? (j < 1 ? j + n : (j > n ? j - n : j)) : j;
}
#define WJ(j) wrap_index(c->crm, (j), c->jx)
#define WI(i) wrap_index(c->crm, (i), c->iy)
#define A3(a, j, i, k) \
  (a)[((size_t)((k)-1) * (size_t)c->iy + (size_t)(WI(i)-1)) * (size_t)c->jx + \
      (size_t)(WJ(j)-1)]
#define A2(a, j, i) (a)[(size_t)(WI(i)-1) * (size_t)c->jx + (size_t)(WJ(j)-1)]
#define QX(n) (c->f->qx + (size_t)((n)-1) * P2(c) * (size_t)c->kz)
#define QXT(n) (c->f->qxten + (size_t)((n)-1) * P2(c) * (size_t)c->kz)
#define TR(n) (c->f->trac + (size_t)((n)-1) * P2(c) * (size_t)c->kz)
#define TRT(n) (c->f->chiten + (size_t)((n)-1) * P2(c) * (size_t)c->kz)

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

static void ctx_free(ctx *c) {
  double **p[] = {&c->gzitak, &c->gzitakh, &c->xkdamp, &c->xknu, &c->zita,
                  &c->zitah, &c->mx2, &c->rmx, &c->rmu, &c->rmv, &c->s,
                  &c->wwkw, &c->tetavf, &c->zdiv2, &c->wz, &c->p0, &c->wfw,
                  &c->wx, &c->zpby, &c->zpbw, &c->ud, &c->vd, &c->laplacian,
                  &c->tkex};
  for (size_t n = 0; n < sizeof(p) / sizeof(p[0]); ++n) {
    free(*p[n]);
    *p[n] = NULL;
  }
}

/* gzita: Share/mod_zita.F90:109-119 */
static double gzita(double zita, double ztop, double a0) {
  double ratio = zita / ztop;
  return (((0.0 - 1.0 * a0) * ratio - (3.0 - 2.0 * a0) * (ratio * ratio)) +
          (2.0 - 1.0 * a0) * (ratio * ratio * ratio)) + 1.0;
}

/* model_zitaf / model_zitah: Share/mod_zita.F90:54-84 */
static void model_zitaf(double *zf, int kz, double ztop) {
  double dz = ztop / (double)kz;
  zf[kz] = 0.0;      /* zitaf(kzp1) */
  zf[0] = ztop;      /* zitaf(1) */
  for (int k = kz; k >= 2; --k) zf[k - 1] = zf[k] + dz;
}
static void model_zitah(double *zh, int kz, double ztop) {
  double dz = ztop / (double)kz;
  zh[kz - 1] = dz * 0.5;
  zh[0] = ztop - dz * 0.5;
  for (int k = kz - 1; k >= 2; --k) zh[k - 1] = zh[k] + dz;
}

static int ctx_init(ctx *c, const orc_grid *g, orc_fields *f) {
  memset(c, 0, sizeof(*c));
  c->g = g;
  c->f = f;
  if (g->jx < 6 || g->iy < 6 || g->kz < 4 || g->nqx < 1 || g->nqx > 10 ||
      g->nadv < 1 || g->nsound < 1 || g->ntr < 0 ||
      (g->ibltyp == 2 && (!f->tke || !f->tketen)) ||
      (g->ntr > 0 && (!f->trac || !f->chiten)) || (g->crm && g->lrotllr))
    return -1;
  c->jx = g->jx; c->iy = g->iy; c->kz = g->kz;
  c->kzp1 = g->kz + 1; c->kzm1 = g->kz - 1;
  c->crm = g->crm != 0;
  if (c->crm) {
    /* Main/mpplib/mod_mppparam.F90:1290-1309: every extent is the whole
     * periodic domain */
    c->jde1 = c->jdi1 = c->jdii1 = c->jce1 = c->jci1 = 1;
    c->jde2 = c->jdi2 = c->jdii2 = c->jce2 = c->jci2 = g->jx;
    c->ide1 = c->idi1 = c->idii1 = c->ice1 = c->ici1 = 1;
    c->ide2 = c->idi2 = c->idii2 = c->ice2 = c->ici2 = g->iy;
    /* init_moloch :288-293 */
    c->jmin = -1; c->jmax = g->jx + 2; c->imin = -1; c->imax = g->iy + 2;
  } else {
    c->jde1 = 1; c->jde2 = g->jx; c->jdi1 = 2; c->jdi2 = g->jx - 1;
    c->jdii1 = 3; c->jdii2 = g->jx - 2;
    c->jce1 = 1; c->jce2 = g->jx - 1; c->jci1 = 2; c->jci2 = g->jx - 2;
    c->ide1 = 1; c->ide2 = g->iy; c->idi1 = 2; c->idi2 = g->iy - 1;
    c->idii1 = 3; c->idii2 = g->iy - 2;
    c->ice1 = 1; c->ice2 = g->iy - 1; c->ici1 = 2; c->ici2 = g->iy - 2;
    /* init_moloch :280-283 (no band / crm) */
    c->jmin = 1; c->jmax = g->jx - 1; c->imin = 1; c->imax = g->iy - 1;
  }
  size_t p2 = P2(c);
  size_t n3 = p2 * (size_t)g->kz, n3p = p2 * (size_t)(g->kz + 1);
  int kz = g->kz;
  c->gzitak = calloc((size_t)kz + 1, sizeof(double));
  c->gzitakh = calloc((size_t)kz, sizeof(double));
  c->xkdamp = calloc((size_t)kz, sizeof(double));
  c->xknu = calloc((size_t)kz, sizeof(double));
  c->zita = calloc((size_t)kz + 1, sizeof(double));
  c->zitah = calloc((size_t)kz, sizeof(double));
  c->mx2 = calloc(p2, sizeof(double));
  c->rmx = calloc(p2, sizeof(double));
  c->rmu = calloc(p2, sizeof(double));
  c->rmv = calloc(p2, sizeof(double));
  c->s = calloc(n3p, sizeof(double));
  c->wwkw = calloc(n3p, sizeof(double));
  c->tetavf = calloc(n3, sizeof(double));
  c->zdiv2 = calloc(n3, sizeof(double));
  c->wz = calloc(n3, sizeof(double));
  c->p0 = calloc(n3, sizeof(double));
  c->wfw = calloc(n3p, sizeof(double));
  c->wx = calloc(n3, sizeof(double));
  c->zpby = calloc(n3, sizeof(double));
  c->zpbw = calloc(n3, sizeof(double));
  c->ud = calloc(n3, sizeof(double));
  c->vd = calloc(n3, sizeof(double));
  c->laplacian = calloc(n3, sizeof(double));
  if (g->ibltyp == 2) {
    c->tkex = calloc(n3, sizeof(double));
    if (!c->tkex) {
      ctx_free(c);
      return -1;
    }
  }
  if (!c->gzitak || !c->gzitakh || !c->xkdamp || !c->xknu || !c->zita ||
      !c->zitah || !c->mx2 || !c->rmx || !c->rmu || !c->rmv || !c->s ||
      !c->wwkw || !c->tetavf || !c->zdiv2 || !c->wz || !c->p0 || !c->wfw ||
      !c->wx || !c->zpby || !c->zpbw || !c->ud || !c->vd || !c->laplacian) {
    ctx_free(c);
    return -1;
  }
  /* vertical grid: Main/mod_params.F90:2460-2463 */
  model_zitaf(c->zita, kz, g->mo_ztop);
  model_zitah(c->zitah, kz, g->mo_ztop);
  double mo_dzita = c->zita[kz - 1];
  c->rdx = 1.0 / g->dx;                          /* mod_params.F90:2199 */
  /* init_moloch: Main/mod_moloch.F90:263-307 */
  for (int i = c->ide1; i <= c->ide2; ++i)
    for (int j = c->jde1; j <= c->jde2; ++j) {
      A2(c->mx2, j, i) = A2(f->mx, j, i) * A2(f->mx, j, i);
      A2(c->rmx, j, i) = 1.0 / A2(f->mx, j, i);
      A2(c->rmu, j, i) = 1.0 / A2(f->mu, j, i);
      A2(c->rmv, j, i) = 1.0 / A2(f->mv, j, i);
    }
  for (int k = 1; k <= kz + 1; ++k)
    c->gzitak[k - 1] = gzita(c->zita[k - 1], g->mo_ztop, g->mo_a0);
  for (int k = 1; k <= kz; ++k)
    c->gzitakh[k - 1] = gzita(c->zitah[k - 1], g->mo_ztop, g->mo_a0);
  c->rdzita = 1.0 / mo_dzita;
  for (int k = 1; k <= kz; ++k) {
    /* numax = 0.125, ddamp = 0.85 (:147-148) */
    c->xkdamp[k - 1] = 0.125 * 0.850 *
        (1.0 / ((double)k + 1.0) - 1.0 / ((double)kz + 2.0));
    c->xknu[k - 1] = 0.125 * (0.55 + 0.45 *
        ((double)(kz - k + 1) - 1.0) / ((double)kz - 1.0));
  }
  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j) A3(f->w, j, i, 1) = 0.0;
  c->dtstepa = g->dtsec / (double)g->nadv;
  c->dtsound = c->dtstepa / (double)g->nsound;
  return 0;
}

/* local_flow_param: Main/mod_moloch.F90:1571-1590.  minnum is the real(4)
 * rounding of 1e-30, widened back to double for the comparison. */
static inline double local_flow_param(double num, double den) {
  const double minden = 1.0e-30;
  const double minnum = (double)(float)1.0e-30;
  if (fabs(den) < minden) return fabs(num) < minnum ? 1.0 : 0.0;
  return num / den;
}

/* superbee-type limiter used at :882, :910, :940, :998, :970, :1026 */
static inline double waf_limiter(double r) {
  return fmaxd(0.0, fmind(2.0, fmaxd(r, fmind(2.0 * r, 1.0))));
}

/* divergence_damping: Main/mod_moloch.F90:738-765 */
static void divergence_damping(ctx *c, double dts) {
  orc_fields *f = c->f;
  double dxrdt = c->g->dx / dts;
  if (c->g->lrotllr) {
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jdi1; j <= c->jdi2; ++j) {
          double xdam = dxrdt * c->xkdamp[k - 1] * A2(f->mu, j, i);
          A3(f->u, j, i, k) = A3(f->u, j, i, k) +
              xdam * (A3(c->zdiv2, j, i, k) - A3(c->zdiv2, j - 1, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->idi1; i <= c->idi2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double xdam = dxrdt * c->xkdamp[k - 1];
          A3(f->v, j, i, k) = A3(f->v, j, i, k) +
              xdam * (A3(c->zdiv2, j, i, k) - A3(c->zdiv2, j, i - 1, k));
        }
  } else {
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jdi1; j <= c->jdi2; ++j) {
          double xdam = dxrdt * c->xkdamp[k - 1] * A2(f->mu, j, i);
          A3(f->u, j, i, k) = A3(f->u, j, i, k) +
              xdam * (A3(c->zdiv2, j, i, k) - A3(c->zdiv2, j - 1, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->idi1; i <= c->idi2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double xdam = dxrdt * c->xkdamp[k - 1] * A2(f->mv, j, i);
          A3(f->v, j, i, k) = A3(f->v, j, i, k) +
              xdam * (A3(c->zdiv2, j, i, k) - A3(c->zdiv2, j, i - 1, k));
        }
  }
}

/* divergence_diffusion: Main/mod_moloch.F90:531-543 */
static void divergence_diffusion(ctx *c) {
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(c->laplacian, j, i, k) =
            (A3(c->zdiv2, j - 1, i, k) + A3(c->zdiv2, j + 1, i, k) +
             A3(c->zdiv2, j, i - 1, k) + A3(c->zdiv2, j, i + 1, k) -
             4.0 * A3(c->zdiv2, j, i, k));
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(c->zdiv2, j, i, k) = A3(c->zdiv2, j, i, k) +
            c->xknu[k - 1] * A3(c->laplacian, j, i, k);
}

/* The elimination of Main/mod_moloch.F90:650-664 for one column of n
 * unknowns w(2..kz) stored at index 0..n-1 (index m <-> level m+2): forward
 * from the last unknown with wwkw above the column = wkup and w above the
 * column = wtop, then back substitution downwards in index with
 * w(1) = wbot.  It solves, for m = 0..n-1,
 *   -zu[m]*w[m-1] + (1 + zu[m] + zd[m])*w[m] - zd[m]*w[m+1] = rhs[m]
 * with w[-1] = wbot and w[n] = wtop (when wkup = 0). */
static void column_solve(int n, const double *zu, const double *zd,
                         const double *rhs, double wtop, double wkup,
                         double wbot, double *w) {
  double wk[n + 1];
  double w_up = wtop, wk_up = wkup;
  for (int m = n - 1; m >= 0; --m) {
    double zrapp = 1.0 / (1.0 + zd[m] + zu[m] - zd[m] * wk_up);
    w[m] = zrapp * (rhs[m] + zd[m] * w_up);
    wk[m] = zrapp * zu[m];
    w_up = w[m];
    wk_up = wk[m];
  }
  double below = wbot;
  for (int m = 0; m < n; ++m) {
    w[m] = w[m] + wk[m] * below;
    below = w[m];
  }
}

/* sound: Main/mod_moloch.F90:545-736 */
static void sound(ctx *c, double dts) {
  orc_fields *f = c->f;
  const int kz = c->kz, kzp1 = c->kzp1;
  double dtrdx = dts * c->rdx;
  double dtrdy = dts * c->rdx;
  double dtrdz = dts * c->rdzita;
  double zcs2 = (dtrdz * dtrdz) * c_rdrcv;

  OMP_FOR
  for (int k = 2; k <= kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(c->tetavf, j, i, k) =
            0.5 * (A3(f->tetav, j, i, k - 1) + A3(f->tetav, j, i, k));

  for (int nsound = 1; nsound <= c->g->nsound; ++nsound) {
    /* :573-578 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k) {
      for (int i = c->ice1; i <= c->ice2; ++i)
        for (int j = c->jde1; j <= c->jde2; ++j)
          A3(c->ud, j, i, k) = A3(f->u, j, i, k);
      for (int i = c->ide1; i <= c->ide2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j)
          A3(c->vd, j, i, k) = A3(f->v, j, i, k);
    }
    /* :582-587 */
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        double zuh = A3(f->u, j, i, kz) * A2(f->hx, j, i) +
                     A3(f->u, j + 1, i, kz) * A2(f->hx, j + 1, i);
        double zvh = A3(f->v, j, i, kz) * A2(f->hy, j, i) +
                     A3(f->v, j, i + 1, kz) * A2(f->hy, j, i + 1);
        A3(c->s, j, i, kzp1) = -0.5 * (zuh + zvh);
        A3(f->w, j, i, kzp1) = -A3(c->s, j, i, kzp1);
      }
    /* :591-597 */
    OMP_FOR
    for (int k = 2; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double zuh = (A3(f->u, j, i, k) + A3(f->u, j, i, k - 1)) *
                           A2(f->hx, j, i) +
                       (A3(f->u, j + 1, i, k) + A3(f->u, j + 1, i, k - 1)) *
                           A2(f->hx, j + 1, i);
          double zvh = (A3(f->v, j, i, k) + A3(f->v, j, i, k - 1)) *
                           A2(f->hy, j, i) +
                       (A3(f->v, j, i + 1, k) + A3(f->v, j, i + 1, k - 1)) *
                           A2(f->hy, j, i + 1);
          A3(c->s, j, i, k) = -0.25 * (zuh + zvh) * c->gzitak[k - 1];
        }
    /* :602-618 */
    if (c->g->lrotllr) {
      OMP_FOR
      for (int k = 1; k <= kz; ++k)
        for (int i = c->ice1; i <= c->ice2; ++i)
          for (int j = c->jce1; j <= c->jce2; ++j) {
            double zum = dtrdx * A3(f->u, j, i, k) * A3(f->rfmzu, j, i, k);
            double zup = dtrdx * A3(f->u, j + 1, i, k) *
                         A3(f->rfmzu, j + 1, i, k);
            double zvm = dtrdy * A3(f->v, j, i, k) * A3(f->rfmzv, j, i, k) *
                         A2(c->rmv, j, i);
            double zvp = dtrdy * A3(f->v, j, i + 1, k) *
                         A3(f->rfmzv, j, i + 1, k) * A2(c->rmv, j, i + 1);
            A3(c->zdiv2, j, i, k) = A3(f->fmz, j, i, k) * A2(f->mx, j, i) *
                                    ((zup - zum) + (zvp - zvm));
          }
    } else {
      OMP_FOR
      for (int k = 1; k <= kz; ++k)
        for (int i = c->ice1; i <= c->ice2; ++i)
          for (int j = c->jce1; j <= c->jce2; ++j) {
            double zum = dtrdx * A3(f->u, j, i, k) * A3(f->rfmzu, j, i, k) *
                         A2(c->rmu, j, i);
            double zup = dtrdx * A3(f->u, j + 1, i, k) *
                         A3(f->rfmzu, j + 1, i, k) * A2(c->rmu, j + 1, i);
            double zvm = dtrdy * A3(f->v, j, i, k) * A3(f->rfmzv, j, i, k) *
                         A2(c->rmv, j, i);
            double zvp = dtrdy * A3(f->v, j, i + 1, k) *
                         A3(f->rfmzv, j, i + 1, k) * A2(c->rmv, j, i + 1);
            A3(c->zdiv2, j, i, k) = A3(f->fmz, j, i, k) * A2(c->mx2, j, i) *
                                    ((zup - zum) + (zvp - zvm));
          }
    }
    if (c->probe_stop == 1) return;
    if (c->g->do_divdamp) divergence_damping(c, dts);
    if (c->g->do_divfilter) divergence_diffusion(c);
    if (c->probe_stop == 2) return;
    /* :627-630 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j)
          A3(c->zdiv2, j, i, k) = A3(c->zdiv2, j, i, k) +
              A3(f->bdywtw, j, i, k) * dtrdz * A3(f->fmz, j, i, k) *
              (A3(c->s, j, i, k) - A3(c->s, j, i, k + 1));
    if (c->probe_stop == 3) return;
    /* :634-664 implicit w: the explicit part and the coefficients of the
     * tridiagonal system per column, then the reference's elimination
     * (column_solve, shared with orc_thomas). */
    OMP_FOR
    for (int i = c->ici1; i <= c->ici2; ++i) {
      double zuv[kz + 1], zdv[kz + 1], rhs[kz + 1], wcol[kz + 1];
      for (int j = c->jci1; j <= c->jci2; ++j) {
        for (int k = kz; k >= 2; --k) {
          A3(c->tetavf, j, i, k) = A3(c->tetavf, j, i, k) -
              A3(f->w, j, i, k) * A3(f->fmzf, j, i, k) * dtrdz *
              (A3(f->tetav, j, i, k - 1) - A3(f->tetav, j, i, k));
          double zrom1w = c_cpd * A3(c->tetavf, j, i, k) * A3(f->fmzf, j, i, k);
          double zwexpl = A3(f->w, j, i, k) - zrom1w * dtrdz *
              (A3(f->pai, j, i, k - 1) - A3(f->pai, j, i, k)) - c_egrav * dts;
          zwexpl = zwexpl + c_rdrcv * zrom1w * dtrdz *
              (A3(f->pai, j, i, k - 1) * A3(c->zdiv2, j, i, k - 1) -
               A3(f->pai, j, i, k) * A3(c->zdiv2, j, i, k));
          zuv[k - 2] = zcs2 * A3(f->fmz, j, i, k - 1) * zrom1w *
                       A3(f->pai, j, i, k - 1) + f->ffilt[k - 1];
          zdv[k - 2] = zcs2 * A3(f->fmz, j, i, k) * zrom1w *
                       A3(f->pai, j, i, k) + f->ffilt[k - 1];
          rhs[k - 2] = zwexpl;
        }
        column_solve(kz - 1, zuv, zdv, rhs, A3(f->w, j, i, kzp1),
                     A3(c->wwkw, j, i, kzp1), A3(f->w, j, i, 1), wcol);
        for (int k = 2; k <= kz; ++k) A3(f->w, j, i, k) = wcol[k - 2];
      }
    }
    /* :668-671 Exner */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j)
          A3(f->pai, j, i, k) = A3(f->pai, j, i, k) *
              (1.0 - c_rdrcv * (A3(c->zdiv2, j, i, k) +
                                (dtrdz * A3(f->fmz, j, i, k) *
                                 (A3(f->w, j, i, k) - A3(f->w, j, i, k + 1)))));
    /* :677-721 horizontal momentum */
    double zfz = c_egrav * dts;
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jdi1; j <= c->jdi2; ++j) {
          double zcx = dtrdx * A2(f->mu, j, i);
          double zrom1u = 0.5 * c_cpd *
              (A3(f->tetav, j - 1, i, k) + A3(f->tetav, j, i, k));
          double zcor1u = A2(f->coru, j, i) * dts * A3(c->vd, j, i, k);
          A3(f->u, j, i, k) = A3(f->u, j, i, k) + A3(f->bdywtu, j, i, k) *
              (zcor1u - zfz * A2(f->hx, j, i) * c->gzitakh[k - 1] -
               zcx * zrom1u * (A3(f->pai, j, i, k) - A3(f->pai, j - 1, i, k)));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->idi1; i <= c->idi2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double zcy = c->g->lrotllr ? dtrdy : dtrdy * A2(f->mv, j, i);
          double zrom1v = 0.5 * c_cpd *
              (A3(f->tetav, j, i - 1, k) + A3(f->tetav, j, i, k));
          double zcor1v = A2(f->corv, j, i) * dts * A3(c->ud, j, i, k);
          A3(f->v, j, i, k) = A3(f->v, j, i, k) + A3(f->bdywtv, j, i, k) *
              (-zcor1v - zfz * A2(f->hy, j, i) * c->gzitakh[k - 1] -
               zcy * zrom1v * (A3(f->pai, j, i, k) - A3(f->pai, j, i - 1, k)));
        }
  }
  /* :728-734 */
  OMP_FOR
  for (int k = 2; k <= kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(c->s, j, i, k) = (A3(f->w, j, i, k) + A3(c->s, j, i, k)) *
                            A3(f->fmzf, j, i, k);
  for (int i = c->ici1; i <= c->ici2; ++i)
    for (int j = c->jci1; j <= c->jci2; ++j) {
      A3(c->s, j, i, 1) = 0.0;
      A3(c->s, j, i, kzp1) = 0.0;
    }
}

/* wafone: Main/mod_moloch.F90:838-1042 */
static void wafone(ctx *c, double *pp, double dta) {
  orc_fields *f = c->f;
  const int kz = c->kz, kzp1 = c->kzp1, kzm1 = c->kzm1;
  double dtrdx = dta * c->rdx;
  double dtrdy = dta * c->rdx;
  double dtrdz = dta * c->rdzita;
  dtrdz = 0.5 * dtrdz;                 /* do_vadvtwice (:130, :858-860) */

  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j) {
      A3(c->wfw, j, i, 1) = 0.0;
      A3(c->wfw, j, i, kzp1) = 0.0;
    }
  /* two vertical half steps: :868-892 then :896-920 */
  for (int pass = 0; pass < 2; ++pass) {
    double *src = pass == 0 ? pp : c->wz;
    OMP_FOR
    for (int k = 1; k <= kzm1; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          double zamu = A3(c->s, j, i, k + 1) * dtrdz;
          double is;
          int k1, k1p1;
          if (zamu >= 0.0) {
            is = 1.0;
            k1 = k + 1;
            k1p1 = k1 + 1;
            if (k1p1 > kz) k1p1 = kz;
          } else {
            is = -1.0;
            k1 = k - 1;
            k1p1 = k;
            if (k1 < 1) k1 = 1;
          }
          double r = local_flow_param(A3(src, j, i, k1) - A3(src, j, i, k1p1),
                                      A3(src, j, i, k) - A3(src, j, i, k + 1));
          double b = waf_limiter(r);
          double zphi = is + zamu * b - is * b;
          A3(c->wfw, j, i, k + 1) = 0.5 * A3(c->s, j, i, k + 1) *
              ((1.0 + zphi) * A3(src, j, i, k + 1) +
               (1.0 - zphi) * A3(src, j, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          double zrfmu = dtrdz * A3(f->fmz, j, i, k) / A3(f->fmzf, j, i, k);
          double zrfmd = dtrdz * A3(f->fmz, j, i, k) / A3(f->fmzf, j, i, k + 1);
          double zdv = (A3(c->s, j, i, k) * zrfmu -
                        A3(c->s, j, i, k + 1) * zrfmd) * A3(src, j, i, k);
          A3(c->wz, j, i, k) = A3(src, j, i, k) -
              A3(c->wfw, j, i, k) * zrfmu + A3(c->wfw, j, i, k + 1) * zrfmd +
              zdv;
        }
  }

  if (c->g->lrotllr) {
    /* :928-953 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ice2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          double zamu = A3(f->v, j, i, k) * dtrdy;
          double is;
          int ih;
          if (zamu > 0.0) { is = 1.0; ih = i - 1; }
          else { is = -1.0; ih = (i + 1 < c->imax) ? i + 1 : c->imax; }
          int ihm1 = (ih - 1 > c->imin) ? ih - 1 : c->imin;
          double r = local_flow_param(A3(c->wz, j, ih, k) - A3(c->wz, j, ihm1, k),
                                      A3(c->wz, j, i, k) - A3(c->wz, j, i - 1, k));
          double b = waf_limiter(r);
          double zphi = is + zamu * b - is * b;
          A3(c->zpby, j, i, k) = 0.5 * A3(f->v, j, i, k) *
              ((1.0 + zphi) * A3(c->wz, j, i - 1, k) +
               (1.0 - zphi) * A3(c->wz, j, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          double zhxvtn = dtrdy * A2(c->rmv, j, i + 1) * A2(f->mx, j, i);
          double zhxvts = dtrdy * A2(c->rmv, j, i) * A2(f->mx, j, i);
          double zrfmn = zhxvtn * A3(f->fmz, j, i, k) * A3(f->rfmzv, j, i + 1, k);
          double zrfms = zhxvts * A3(f->fmz, j, i, k) * A3(f->rfmzv, j, i, k);
          double zdv = (A3(f->v, j, i + 1, k) * zrfmn -
                        A3(f->v, j, i, k) * zrfms) * A3(pp, j, i, k);
          A3(c->p0, j, i, k) = A3(c->wz, j, i, k) +
              (A3(c->zpby, j, i, k) * zrfms - A3(c->zpby, j, i + 1, k) * zrfmn +
               zdv);
        }
    /* :959-982 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jce2; ++j) {
          double zamu = A3(f->u, j, i, k) * A2(f->mu, j, i) * dtrdx;
          double is;
          int jh;
          if (zamu > 0.0) { is = 1.0; jh = j - 1; }
          else { is = -1.0; jh = (j + 1 < c->jmax) ? j + 1 : c->jmax; }
          int jhm1 = (jh - 1 > c->jmin) ? jh - 1 : c->jmin;
          double r = local_flow_param(A3(c->p0, jh, i, k) - A3(c->p0, jhm1, i, k),
                                      A3(c->p0, j, i, k) - A3(c->p0, j - 1, i, k));
          double b = waf_limiter(r);
          double zphi = is + zamu * b - is * b;
          A3(c->zpbw, j, i, k) = 0.5 * A3(f->u, j, i, k) *
              ((1.0 + zphi) * A3(c->p0, j - 1, i, k) +
               (1.0 - zphi) * A3(c->p0, j, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double zcostx = dtrdx * A2(f->mx, j, i);
          double zrfmw = zcostx * A3(f->fmz, j, i, k) * A3(f->rfmzu, j, i, k);
          double zrfme = zcostx * A3(f->fmz, j, i, k) * A3(f->rfmzu, j + 1, i, k);
          double zdv = (A3(f->u, j + 1, i, k) * zrfme -
                        A3(f->u, j, i, k) * zrfmw) * A3(pp, j, i, k);
          A3(pp, j, i, k) = A3(c->p0, j, i, k) +
              A3(c->zpbw, j, i, k) * zrfmw - A3(c->zpbw, j + 1, i, k) * zrfme +
              zdv;
        }
  } else {
    /* :986-1010 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ice2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          double zamu = A3(f->v, j, i, k) * A2(f->mv, j, i) * dtrdy;
          double is;
          int ih;
          if (zamu > 0.0) { is = 1.0; ih = i - 1; }
          else { is = -1.0; ih = (i + 1 < c->imax) ? i + 1 : c->imax; }
          int ihm1 = (ih - 1 > c->imin) ? ih - 1 : c->imin;
          double r = local_flow_param(A3(c->wz, j, ih, k) - A3(c->wz, j, ihm1, k),
                                      A3(c->wz, j, i, k) - A3(c->wz, j, i - 1, k));
          double b = waf_limiter(r);
          double zphi = is + zamu * b - is * b;
          A3(c->zpby, j, i, k) = 0.5 * A3(f->v, j, i, k) *
              ((1.0 + zphi) * A3(c->wz, j, i - 1, k) +
               (1.0 - zphi) * A3(c->wz, j, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jce1; j <= c->jce2; ++j) {
          /* rfmzu, not rfmzv, on the meridional fluxes (:1004-1005) */
          double zrfmn = dtrdy * A3(f->fmz, j, i, k) * A3(f->rfmzu, j, i + 1, k);
          double zrfms = dtrdy * A3(f->fmz, j, i, k) * A3(f->rfmzu, j, i, k);
          double zdv = (A3(f->v, j, i + 1, k) * A2(c->rmv, j, i + 1) * zrfmn -
                        A3(f->v, j, i, k) * A2(c->rmv, j, i) * zrfms) *
                       A3(pp, j, i, k);
          A3(c->p0, j, i, k) = A3(c->wz, j, i, k) + A2(c->mx2, j, i) *
              (A3(c->zpby, j, i, k) * zrfms - A3(c->zpby, j, i + 1, k) * zrfmn +
               zdv);
        }
    /* :1015-1038 */
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jce2; ++j) {
          double zamu = A3(f->u, j, i, k) * A2(f->mu, j, i) * dtrdx;
          double is;
          int jh;
          if (zamu > 0.0) { is = 1.0; jh = j - 1; }
          else { is = -1.0; jh = (j + 1 < c->jmax) ? j + 1 : c->jmax; }
          int jhm1 = (jh - 1 > c->jmin) ? jh - 1 : c->jmin;
          double r = local_flow_param(A3(c->p0, jh, i, k) - A3(c->p0, jhm1, i, k),
                                      A3(c->p0, j, i, k) - A3(c->p0, j - 1, i, k));
          double b = waf_limiter(r);
          double zphi = is + zamu * b - is * b;
          A3(c->zpbw, j, i, k) = 0.5 * A3(f->u, j, i, k) *
              ((1.0 + zphi) * A3(c->p0, j - 1, i, k) +
               (1.0 - zphi) * A3(c->p0, j, i, k));
        }
    OMP_FOR
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double zrfmw = dtrdx * A3(f->fmz, j, i, k) * A3(f->rfmzu, j, i, k);
          double zrfme = dtrdx * A3(f->fmz, j, i, k) * A3(f->rfmzu, j + 1, i, k);
          double zdv = (A3(f->u, j + 1, i, k) * A2(c->rmu, j + 1, i) * zrfme -
                        A3(f->u, j, i, k) * A2(c->rmu, j, i) * zrfmw) *
                       A3(pp, j, i, k);
          A3(pp, j, i, k) = A3(c->p0, j, i, k) + A2(c->mx2, j, i) *
              (A3(c->zpbw, j, i, k) * zrfmw - A3(c->zpbw, j + 1, i, k) * zrfme +
               zdv);
        }
  }
}

/* zstagtoh: Main/mod_moloch.F90:1445-1459 */
static void zstagtoh(ctx *c, const double *fl, double *hl) {
  const int kz = c->kz;
  OMP_FOR
  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j) {
      for (int k = 2; k <= kz - 1; ++k)
        A3(hl, j, i, k) = 0.5625 * (A3(fl, j, i, k + 1) + A3(fl, j, i, k)) -
                          0.0625 * (A3(fl, j, i, k + 2) + A3(fl, j, i, k - 1));
      A3(hl, j, i, 1) = 0.5 * (A3(fl, j, i, 2) + A3(fl, j, i, 1));
      A3(hl, j, i, kz) = 0.5 * (A3(fl, j, i, kz + 1) + A3(fl, j, i, kz));
    }
}

/* htozstag: Main/mod_moloch.F90:1461-1475 */
static void htozstag(ctx *c, const double *hl, double *fl) {
  const int kz = c->kz;
  OMP_FOR
  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j) {
      for (int k = 3; k <= kz - 1; ++k)
        A3(fl, j, i, k) = 0.5625 * (A3(hl, j, i, k) + A3(hl, j, i, k - 1)) -
                          0.0625 * (A3(hl, j, i, k + 1) + A3(hl, j, i, k - 2));
      A3(fl, j, i, 2) = 0.5 * (A3(hl, j, i, 2) + A3(hl, j, i, 1));
      A3(fl, j, i, kz) = 0.5 * (A3(hl, j, i, kz) + A3(hl, j, i, kz - 1));
    }
}

/* uvxtouvstag: Main/mod_moloch.F90:1477-1522 */
static void uvxtouvstag(ctx *c) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k) {
    for (int i = c->ici1; i <= c->ici2; ++i) {
      for (int j = c->jdii1; j <= c->jdii2; ++j)
        A3(f->u, j, i, k) = 0.5625 * (A3(f->ux, j, i, k) + A3(f->ux, j - 1, i, k)) -
                            0.0625 * (A3(f->ux, j + 1, i, k) + A3(f->ux, j - 2, i, k));
      if (c->crm) continue;                      /* has_bdyright/left */
      A3(f->u, c->jdi2, i, k) =
          0.5 * (A3(f->ux, c->jci2, i, k) + A3(f->ux, c->jce2, i, k));
      A3(f->u, c->jdi1, i, k) =
          0.5 * (A3(f->ux, c->jci1, i, k) + A3(f->ux, c->jce1, i, k));
    }
    for (int i = c->idii1; i <= c->idii2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(f->v, j, i, k) = 0.5625 * (A3(f->vx, j, i, k) + A3(f->vx, j, i - 1, k)) -
                            0.0625 * (A3(f->vx, j, i + 1, k) + A3(f->vx, j, i - 2, k));
    for (int j = c->jci1; j <= c->jci2 && !c->crm; ++j) {   /* has_bdytop/bottom */
      A3(f->v, j, c->idi2, k) =
          0.5 * (A3(f->vx, j, c->ici2, k) + A3(f->vx, j, c->ice2, k));
      A3(f->v, j, c->idi1, k) =
          0.5 * (A3(f->vx, j, c->ici1, k) + A3(f->vx, j, c->ice1, k));
    }
  }
}

/* uvstagtouvx: Main/mod_moloch.F90:1524-1569 */
static void uvstagtouvx(ctx *c) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k) {
    for (int i = c->ice1; i <= c->ice2; ++i) {
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(f->ux, j, i, k) = 0.5625 * (A3(f->u, j + 1, i, k) + A3(f->u, j, i, k)) -
                             0.0625 * (A3(f->u, j + 2, i, k) + A3(f->u, j - 1, i, k));
      if (c->crm) continue;                      /* has_bdyleft/right */
      A3(f->ux, c->jce1, i, k) =
          0.5 * (A3(f->u, c->jde1, i, k) + A3(f->u, c->jdi1, i, k));
      A3(f->ux, c->jce2, i, k) =
          0.5 * (A3(f->u, c->jde2, i, k) + A3(f->u, c->jdi2, i, k));
    }
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(f->vx, j, i, k) = 0.5625 * (A3(f->v, j, i + 1, k) + A3(f->v, j, i, k)) -
                             0.0625 * (A3(f->v, j, i + 2, k) + A3(f->v, j, i - 1, k));
    for (int j = c->jce1; j <= c->jce2 && !c->crm; ++j) {   /* has_bdybottom/top */
      A3(f->vx, j, c->ice1, k) =
          0.5 * (A3(f->v, j, c->ide1, k) + A3(f->v, j, c->idi1, k));
      A3(f->vx, j, c->ice2, k) =
          0.5 * (A3(f->v, j, c->ide2, k) + A3(f->v, j, c->idi2, k));
    }
  }
}

/* advection: Main/mod_moloch.F90:767-836 */
/* curvature terms of advection: Main/mod_moloch.F90:811-825.  vx is
 * updated with the already-updated ux, as the reference's statement order
 * gives. */
static void curvature(ctx *c, double dta) {
  orc_fields *f = c->f;
  if (c->g->lrotllr) {
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double dlat = c_degrad * 0.5 * (f->rlat[i - 1] + f->rlat[i]);
          double tanx = sin(dlat) * A2(f->mx, j, i) * c_rearthrad;
          A3(f->ux, j, i, k) = A3(f->ux, j, i, k) +
              A3(f->ux, j, i, k) * A3(f->vx, j, i, k) * tanx * dta;
          A3(f->vx, j, i, k) = A3(f->vx, j, i, k) -
              A3(f->ux, j, i, k) * A3(f->ux, j, i, k) * tanx * dta;
        }
  } else {
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          double tanx = (A2(f->mu, j - 1, i) - A2(f->mu, j, i)) * c->rdx;
          double tany = (A2(f->mv, j, i - 1) - A2(f->mv, j, i)) * c->rdx;
          A3(f->ux, j, i, k) = A3(f->ux, j, i, k) +
              A3(f->ux, j, i, k) * A3(f->vx, j, i, k) * tanx * dta;
          A3(f->vx, j, i, k) = A3(f->vx, j, i, k) -
              A3(f->ux, j, i, k) * A3(f->ux, j, i, k) * tany * dta;
        }
  }
}

static void advection(ctx *c, double dta) {
  orc_fields *f = c->f;
  uvstagtouvx(c);
  zstagtoh(c, f->w, c->wx);
  if (c->g->ibltyp == 2) zstagtoh(c, f->tke, c->tkex);         /* :782-784 */
  wafone(c, f->tetav, dta);
  wafone(c, f->pai, dta);
  wafone(c, f->ux, dta);
  wafone(c, f->vx, dta);
  wafone(c, c->wx, dta);
  wafone(c, QX(1), dta);
  if (c->g->ipptls > 0)
    for (int n = 2; n <= c->g->nqx; ++n) wafone(c, QX(n), dta);
  if (c->g->ibltyp == 2) wafone(c, c->tkex, dta);              /* :799-801 */
  for (int n = 1; n <= c->g->ntr; ++n) wafone(c, TR(n), dta);  /* :802-807 */
  curvature(c, dta);
  uvxtouvstag(c);
  htozstag(c, c->wx, f->w);
  if (c->g->ibltyp == 2) htozstag(c, c->tkex, f->tke);         /* :831-834 */
}

/* Tv denominator / factor: temp_to_tvirt :1608-1627, tvirt_to_temp :1629-1648 */
static inline double tv_factor(ctx *c, size_t idx) {
  const size_t b = P2(c) * (size_t)c->kz;
  const double *q = c->f->qx;
  if (c->g->ipptls > 1)
    return 1.0 + c_ep1 * q[idx] - q[b + idx] - q[2 * b + idx] - q[3 * b + idx] -
           q[4 * b + idx];
  if (c->g->ipptls > 0) return 1.0 + c_ep1 * q[idx] - q[b + idx];
  return 1.0 + c_ep1 * q[idx];
}

static void tvirt_to_temp(ctx *c) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        size_t idx = &A3(f->t, j, i, k) - f->t;
        f->t[idx] = f->tvirt[idx] / tv_factor(c, idx);
      }
}

static void temp_to_tvirt(ctx *c) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) {
        size_t idx = &A3(f->t, j, i, k) - f->t;
        f->tvirt[idx] = f->t[idx] * tv_factor(c, idx);
      }
}

/* reset_tendencies: Main/mod_moloch.F90:1044-1083 (cldfra/cldlwc belong to
 * physics and are not held here) */
static void reset_tendencies(ctx *c) {
  orc_fields *f = c->f;
  for (int k = 1; k <= c->kzp1; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) A3(c->s, j, i, k) = 0.0;
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) A3(c->zdiv2, j, i, k) = 0.0;
  for (int k = 2; k <= c->kzp1; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) A3(c->wwkw, j, i, k) = 0.0;
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        A3(f->tten, j, i, k) = 0.0;
        A3(f->uten, j, i, k) = 0.0;
        A3(f->vten, j, i, k) = 0.0;
        for (int n = 1; n <= c->g->nqx; ++n) A3(QXT(n), j, i, k) = 0.0;
        for (int n = 1; n <= c->g->ntr; ++n) A3(TRT(n), j, i, k) = 0.0;   /* :1066-1070 */
      }
  if (c->g->ibltyp == 2)                                        /* :1071-1075 */
    for (int k = 1; k <= c->kzp1; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) A3(f->tketen, j, i, k) = 0.0;
}

/* dynamical_core: Main/mod_moloch.F90:1085-1141 (idiag = 0, ichem = 0) */
static void dynamical_core(ctx *c) {
  orc_fields *f = c->f;
  for (int nadv = 1; nadv <= c->g->nadv; ++nadv) {
    sound(c, c->dtsound);
    advection(c, c->dtstepa);
  }
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(f->tvirt, j, i, k) = A3(f->tetav, j, i, k) * A3(f->pai, j, i, k);
  tvirt_to_temp(c);
}

/* pfwsat: Share/pfwsat.inc:2-43 */
double orc_pfwsat(double t, double p) {
  const double a0 = 0.611213476e+03, a1 = 0.444007856e+02,
               a2 = 0.143064234e+01, a3 = 0.264461437e-01,
               a4 = 0.305903558e-03, a5 = 0.196237241e-05,
               a6 = 0.892344772e-08, a7 = -0.373208410e-10,
               a8 = 0.209339997e-13;
  const double c0 = 0.611123516e+03, c1 = 0.503109514e+02,
               c2 = 0.188369801e+01, c3 = 0.420547422e-01,
               c4 = 0.614396778e-03, c5 = 0.602780717e-05,
               c6 = 0.387940929e-07, c7 = 0.149436277e-09,
               c8 = 0.262655803e-12;
  double t_limit = t - c_tzero;
  double td = fmind(fmaxd(t_limit, -75.0), 100.0);
  double es;
  if (td >= 0.0)
    es = fmind(a0 + td * (a1 + td * (a2 + td * (a3 + td * (a4 + td * (a5 +
                td * (a6 + td * (a7 + td * a8))))))), 0.15 * p);
  else
    es = fmind(c0 + td * (c1 + td * (c2 + td * (c3 + td * (c4 + td * (c5 +
                td * (c6 + td * (c7 + td * c8))))))), 0.15 * p);
  return c_ep2 * (es / (p - es));
}

/* moloch :348-352 diagnostics */
static void diagnostics(ctx *c) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) {
        size_t idx = &A3(f->t, j, i, k) - f->t;
        f->p[idx] = pow(f->pai[idx], c_cpovr) * c_p00;
        f->rho[idx] = f->p[idx] / (c_rgas * f->t[idx]);
        f->qsat[idx] = orc_pfwsat(f->t[idx], f->p[idx]);
      }
}

/* extrapolate_surface_pressure: Main/mod_moloch.F90:1592-1606 */
static void extrapolate_surface_pressure(ctx *c) {
  orc_fields *f = c->f;
  const int kz = c->kz, kzm1 = c->kzm1;
  for (int i = c->ici1; i <= c->ici2; ++i)
    for (int j = c->jci1; j <= c->jci2; ++j) {
      double lrt = (A3(f->tvirt, j, i, kzm1) - A3(f->tvirt, j, i, kz)) /
                   (A3(f->zeta, j, i, kzm1) - A3(f->zeta, j, i, kz));
      if (lrt < -c_govcp)
        lrt = -c_govcp;
      else if (lrt > -0.005)
        lrt = 0.65 * lrt - 0.35 * c_lrate;
      double tv = A3(f->tvirt, j, i, kz) - lrt * 0.5 * A3(f->zeta, j, i, kz);
      A2(f->ps, j, i) = A3(f->p, j, i, kz) *
                        exp(c_govr * A3(f->zeta, j, i, kz) / tv);
    }
}

/* status_update: Main/mod_moloch.F90:1403-1443 */
static void status_update(ctx *c, double dtinc) {
  orc_fields *f = c->f;
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        A3(f->t, j, i, k) = A3(f->t, j, i, k) + dtinc * A3(f->tten, j, i, k);
        A3(f->ux, j, i, k) = A3(f->ux, j, i, k) + dtinc * A3(f->uten, j, i, k);
        A3(f->vx, j, i, k) = A3(f->vx, j, i, k) + dtinc * A3(f->vten, j, i, k);
      }
  for (int n = 1; n <= c->g->nqx; ++n) {
    double *q = QX(n), *qt = QXT(n);
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          A3(q, j, i, k) = A3(q, j, i, k) + dtinc * A3(qt, j, i, k);
          if (A3(q, j, i, k) < qxcheckval[n - 1])
            A3(q, j, i, k) = qxzeroval[n - 1];
        }
  }
  if (c->g->ibltyp == 2) {                                      /* :1419-1424 */
    OMP_FOR
    for (int k = 1; k <= c->kzp1; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          A3(f->tke, j, i, k) = A3(f->tke, j, i, k) + dtinc * A3(f->tketen, j, i, k);
          if (A3(f->tke, j, i, k) < c->g->tkemin) A3(f->tke, j, i, k) = c->g->tkemin;
        }
  }
  for (int n = 1; n <= c->g->ntr; ++n) {                        /* :1425-1430 */
    double *q = TR(n), *qt = TRT(n);
    OMP_FOR
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j) {
          A3(q, j, i, k) = A3(q, j, i, k) + dtinc * A3(qt, j, i, k);
          if (A3(q, j, i, k) < 0.0) A3(q, j, i, k) = 0.0;
        }
  }
  temp_to_tvirt(c);
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) {
        size_t idx = &A3(f->t, j, i, k) - f->t;
        f->tetav[idx] = f->tvirt[idx] / f->pai[idx];
        f->rho[idx] = f->p[idx] / (c_rgas * f->t[idx]);
        f->qsat[idx] = orc_pfwsat(f->t[idx], f->p[idx]);
      }
  uvxtouvstag(c);
}

/* ---- lateral boundaries ------------------------------------------------ */
#define BQ(a, j, i, k) A3(a, j, i, k)
#define LIN(b0, b1, j, i, k) (x0 * A3(b0, j, i, k) + x1 * A3(b1, j, i, k))

/* bdyval, MOLOCH branch: Main/mod_bdycod.F90:1638-1866 (single domain: every
 * has_bdy* flag is true) */
static void bdyval(ctx *c, const orc_bdy *b) {
  orc_fields *f = c->f;
  const double x1 = b->x1;
  const double x0 = 1.0 - x1;
  const int kz = c->kz, nqx = c->g->nqx;
  const int pqc = b->present_qc, pqi = b->present_qi;
  const int iqv = 1, iqc = 2, iqi = 3, iqfrst = 2;
  /* west (:1642-1697), then east (:1701-1756): corners excluded */
  for (int side = 0; side < 2; ++side) {
    const int jd = side == 0 ? c->jde1 : c->jde2;   /* u column */
    const int jc = side == 0 ? c->jce1 : c->jce2;   /* cross column */
    const int jn = side == 0 ? c->jci1 : c->jci2;   /* interior neighbour */
    for (int i = c->ici1; i <= c->ici2; ++i)
      A2(f->ps, jc, i) = x0 * A2(b->ps0, jc, i) + x1 * A2(b->ps1, jc, i);
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int k = 1; k <= kz; ++k) A3(f->u, jd, i, k) = LIN(b->u0, b->u1, jd, i, k);
    for (int i = c->idi1; i <= c->idi2; ++i)
      for (int k = 1; k <= kz; ++k) A3(f->v, jc, i, k) = LIN(b->v0, b->v1, jc, i, k);
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int k = 1; k <= kz; ++k) {
        A3(f->t, jc, i, k) = LIN(b->t0, b->t1, jc, i, k);
        A3(f->pai, jc, i, k) = LIN(b->pai0, b->pai1, jc, i, k);
        A3(QX(iqv), jc, i, k) = LIN(b->qv0, b->qv1, jc, i, k);
      }
    if (pqc)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int k = 1; k <= kz; ++k) A3(QX(iqc), jc, i, k) = LIN(b->qc0, b->qc1, jc, i, k);
    if (pqi && c->g->ipptls > 1)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int k = 1; k <= kz; ++k) A3(QX(iqi), jc, i, k) = LIN(b->qi0, b->qi1, jc, i, k);
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int k = 1; k <= kz; ++k)
        for (int n = iqfrst; n <= nqx; ++n) {
          if (pqc && n == iqc) continue;
          if (pqi && n == iqi) continue;
          const double qxint = A3(QX(n), jn, i, k);
          const double uu = A3(f->u, jd, i, k);
          const int inflow = side == 0 ? uu > 0.0 : uu < 0.0;
          A3(QX(n), jc, i, k) = inflow ? qxzeroval[n - 1] : qxint;
        }
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int k = 1; k <= kz; ++k) {
        const double uu = A3(f->u, jd, i, k);
        const int inflow = side == 0 ? uu > 0.0 : uu < 0.0;
        A3(f->w, jc, i, k) = inflow ? 0.0 : A3(f->w, jn, i, k);
      }
    if (c->g->ibltyp == 2) {                         /* :1684-1694, :1742-1752 */
      for (int i = c->ici1; i <= c->ici2; ++i) A3(f->tke, jc, i, 1) = c->g->tkemin;
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int k = 2; k <= kz; ++k) {
          const double us = A3(f->u, jd, i, k) + A3(f->u, jd, i, k - 1);
          const int inflow = side == 0 ? us > 0.0 : us < 0.0;
          A3(f->tke, jc, i, k + 1) = inflow ? c->g->tkemin : A3(f->tke, jn, i, k + 1);
        }
    }
  }
  /* south (:1760-1813), then north (:1817-1866): corners included */
  for (int side = 0; side < 2; ++side) {
    const int id = side == 0 ? c->ide1 : c->ide2;
    const int ic = side == 0 ? c->ice1 : c->ice2;
    const int in = side == 0 ? c->ici1 : c->ici2;
    for (int j = c->jce1; j <= c->jce2; ++j)
      A2(f->ps, j, ic) = x0 * A2(b->ps0, j, ic) + x1 * A2(b->ps1, j, ic);
    for (int j = c->jde1; j <= c->jde2; ++j)
      for (int k = 1; k <= kz; ++k) A3(f->u, j, ic, k) = LIN(b->u0, b->u1, j, ic, k);
    for (int j = c->jce1; j <= c->jce2; ++j)
      for (int k = 1; k <= kz; ++k) A3(f->v, j, id, k) = LIN(b->v0, b->v1, j, id, k);
    for (int j = c->jce1; j <= c->jce2; ++j)
      for (int k = 1; k <= kz; ++k) {
        A3(f->t, j, ic, k) = LIN(b->t0, b->t1, j, ic, k);
        A3(f->pai, j, ic, k) = LIN(b->pai0, b->pai1, j, ic, k);
        A3(QX(iqv), j, ic, k) = LIN(b->qv0, b->qv1, j, ic, k);
      }
    if (pqc)
      for (int j = c->jce1; j <= c->jce2; ++j)
        for (int k = 1; k <= kz; ++k) A3(QX(iqc), j, ic, k) = LIN(b->qc0, b->qc1, j, ic, k);
    if (pqi && c->g->ipptls > 1)
      for (int j = c->jce1; j <= c->jce2; ++j)
        for (int k = 1; k <= kz; ++k) A3(QX(iqi), j, ic, k) = LIN(b->qi0, b->qi1, j, ic, k);
    for (int j = c->jce1; j <= c->jce2; ++j)
      for (int k = 1; k <= kz; ++k)
        for (int n = iqfrst; n <= nqx; ++n) {
          if (pqc && n == iqc) continue;
          if (pqi && n == iqi) continue;
          const double qxint = A3(QX(n), j, in, k);
          const double vv = A3(f->v, j, id, k);
          const int inflow = side == 0 ? vv > 0.0 : vv < 0.0;
          A3(QX(n), j, ic, k) = inflow ? qxzeroval[n - 1] : qxint;
        }
    for (int j = c->jce1; j <= c->jce2; ++j)
      for (int k = 1; k <= kz; ++k) {
        const double vv = A3(f->v, j, id, k);
        const int inflow = side == 0 ? vv > 0.0 : vv < 0.0;
        A3(f->w, j, ic, k) = inflow ? 0.0 : A3(f->w, j, in, k);
      }
    if (c->g->ibltyp == 2) {                         /* :1800-1810, :1858-1868 */
      for (int j = c->jce1; j <= c->jce2; ++j) A3(f->tke, j, ic, 1) = c->g->tkemin;
      for (int j = c->jce1; j <= c->jce2; ++j)
        for (int k = 2; k <= kz; ++k) {
          const double vs = A3(f->v, j, id, k) + A3(f->v, j, id, k - 1);
          const int inflow = side == 0 ? vs > 0.0 : vs < 0.0;
          A3(f->tke, j, ic, k + 1) = inflow ? c->g->tkemin : A3(f->tke, j, in, k + 1);
        }
    }
  }
}

/* motopnudge(t,u,v,w,0,dtsec): Main/mod_bdycod.F90:4049-4081 */
static void motopnudge(ctx *c, const orc_bdy *b, double dta) {
  orc_fields *f = c->f;
  const double trtau = 1.0 / (2.0 * 3600.0);
  const double wrtau = 1.0 / (3600.0 / 4.0);
  const double wfac = 0.0;
  const double x1 = b->x1;
  const double x0 = 1.0 - x1;
  for (int i = c->ici1; i <= c->ici2; ++i)
    for (int j = c->jci1; j <= c->jci2; ++j) {
      const double fext = wfac * A3(f->w, j, i, 2);
      const double xf = wrtau * dta;
      A3(f->w, j, i, 2) = (1.0 - xf) * A3(f->w, j, i, 2) + xf * fext;
    }
  for (int k = 1; k <= b->nztop; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        const double fext = LIN(b->t0, b->t1, j, i, k);
        const double xf = b->tnudge[k - 1] * trtau * dta;
        A3(f->t, j, i, k) = (1.0 - xf) * A3(f->t, j, i, k) + xf * fext;
      }
  for (int k = 1; k <= b->nztop; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jdi1; j <= c->jdi2; ++j) {
        const double fext = LIN(b->u0, b->u1, j, i, k);
        const double xf = b->tnudge[k - 1] * trtau * dta;
        A3(f->u, j, i, k) = (1.0 - xf) * A3(f->u, j, i, k) + xf * fext;
      }
  for (int k = 1; k <= b->nztop; ++k)
    for (int i = c->idi1; i <= c->idi2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        const double fext = LIN(b->v0, b->v1, j, i, k);
        const double xf = b->tnudge[k - 1] * trtau * dta;
        A3(f->v, j, i, k) = (1.0 - xf) * A3(f->v, j, i, k) + xf * fext;
      }
}

/* morelax_external (:3996-4031) over j1..j2 x i1..i2 with weights xf */
static void morelax_external(ctx *c, int j1, int j2, int i1, int i2, const double *xfw,
                             double *fld, const double *b0, const double *b1, double x1) {
  const double x0 = 1.0 - x1;
  for (int k = 1; k <= c->kz; ++k)
    for (int i = i1; i <= i2; ++i)
      for (int j = j1; j <= j2; ++j) {
        const double xf = A3(xfw, j, i, k);
        if (xf >= 0.0) {     /* ib > 0 */
          const double fext = (x0 * A3(b0, j, i, k) + x1 * A3(b1, j, i, k));
          A3(fld, j, i, k) = (1.0 - xf) * A3(fld, j, i, k) + xf * fext;
        }
      }
}

/* morelax_fraction (:3962-3994) */
static void morelax_fraction(ctx *c, int j1, int j2, int i1, int i2, const double *xfw,
                             double *fld, double frac) {
  for (int k = 1; k <= c->kz; ++k)
    for (int i = i1; i <= i2; ++i)
      for (int j = j1; j <= j2; ++j) {
        const double xf = A3(xfw, j, i, k);
        if (xf >= 0.0)
          A3(fld, j, i, k) = (1.0 - xf) * A3(fld, j, i, k) + xf * A3(fld, j, i, k) * frac;
      }
}

/* lowpass_filter (Main/mod_bdycod.F90:3898-3938) on one level of zn (a
 * jx x iy plane), single domain (row_reduce / column_reduce are identities) */
static void lowpass_filter(ctx *c, const orc_bdy *b, int j1, int j2, int i1, int i2,
                           int jj1, int jj2, int ii1, int ii2, double *zn,
                           double *sx, double *sy) {
  const int jx = c->jx, iy = c->iy, km2 = 2 * b->km, lm2 = 2 * b->lm;
  const size_t np = (size_t)jx * (size_t)iy;
#define ZN(j, i) zn[(size_t)((i)-1) * (size_t)jx + (size_t)((j)-1)]
#define BVX(j, m) b->bvx[(size_t)((m)-1) * (size_t)jx + (size_t)((j)-1)]
#define BVY(i, l) b->bvy[(size_t)((l)-1) * (size_t)iy + (size_t)((i)-1)]
  for (int k = 1; k <= km2; ++k)
    for (int i = i1; i <= i2; ++i) {
      double acc = 0.0;
      for (int j = jj1; j <= jj2; ++j) acc = acc + ZN(j, i) * BVX(j, k);
      sx[(size_t)(k - 1) * iy + (i - 1)] = acc;
    }
  for (size_t n = 0; n < np; ++n) zn[n] = 0.0;
  for (int k = 1; k <= km2; ++k)
    for (int i = i1; i <= i2; ++i)
      for (int j = j1; j <= j2; ++j)
        ZN(j, i) = ZN(j, i) + sx[(size_t)(k - 1) * iy + (i - 1)] * BVX(j, k);
  for (int l = 1; l <= lm2; ++l)
    for (int j = j1; j <= j2; ++j) {
      double acc = 0.0;
      for (int i = ii1; i <= ii2; ++i) acc = acc + ZN(j, i) * BVY(i, l);
      sy[(size_t)(l - 1) * jx + (j - 1)] = acc;
    }
  for (size_t n = 0; n < np; ++n) zn[n] = 0.0;
  for (int l = 1; l <= lm2; ++l)
    for (int i = i1; i <= i2; ++i)
      for (int j = j1; j <= j2; ++j)
        ZN(j, i) = ZN(j, i) + sy[(size_t)(l - 1) * jx + (j - 1)] * BVY(i, l);
#undef ZN
#undef BVX
#undef BVY
}

/* mospectral_nudge (Main/mod_bdycod.F90:3940-3960) */
static void mospectral_nudge(ctx *c, const orc_bdy *b, int j1, int j2, int i1, int i2,
                             int jj1, int jj2, int ii1, int ii2, double *fld,
                             const double *b0, const double *b1) {
  const double x1 = b->x1;
  const double x0 = 1.0 - x1;
  const size_t np = (size_t)c->jx * (size_t)c->iy;
  double *zn = calloc(np, sizeof(double));
  double *sx = calloc((size_t)c->iy * (size_t)(2 * b->km), sizeof(double));
  double *sy = calloc((size_t)c->jx * (size_t)(2 * b->lm), sizeof(double));
  if (!zn || !sx || !sy) {
    free(zn); free(sx); free(sy);
    return;
  }
  for (int k = 1; k <= c->kz; ++k) {
    for (int i = i1; i <= i2; ++i)
      for (int j = j1; j <= j2; ++j)
        zn[(size_t)(i - 1) * c->jx + (j - 1)] =
            (x0 * A3(b0, j, i, k) + x1 * A3(b1, j, i, k)) - A3(fld, j, i, k);
    lowpass_filter(c, b, j1, j2, i1, i2, jj1, jj2, ii1, ii2, zn, sx, sy);
    for (int i = ii1; i <= ii2; ++i)
      for (int j = jj1; j <= jj2; ++j)
        A3(fld, j, i, k) = A3(fld, j, i, k) +
                           b->cnudge[k - 1] * zn[(size_t)(i - 1) * c->jx + (j - 1)];
  }
  free(zn); free(sx); free(sy);
}

/* boundary: Main/mod_moloch.F90:448-529 (idiag = 0, ichem = 0) */
static void boundary(ctx *c, orc_bdy *b) {
  orc_fields *f = c->f;
  bdyval(c, b);
  if (b->top_nudge) motopnudge(c, b, c->g->dtsec);
  morelax_external(c, c->jdi1, c->jdi2, c->ici1, c->ici2, b->xfu, f->u, b->u0, b->u1, b->x1);
  morelax_external(c, c->jci1, c->jci2, c->idi1, c->idi2, b->xfv, f->v, b->v0, b->v1, b->x1);
  morelax_external(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, f->t, b->t0, b->t1, b->x1);
  morelax_external(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, f->pai, b->pai0, b->pai1,
                   b->x1);
  morelax_external(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, QX(1), b->qv0, b->qv1, b->x1);
  morelax_fraction(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, f->w, 0.0);
  if (c->g->ipptls > 0) {
    if (b->present_qc)
      morelax_external(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, QX(2), b->qc0, b->qc1,
                       b->x1);
    if (c->g->ipptls > 1 && b->present_qi)
      morelax_external(c, c->jci1, c->jci2, c->ici1, c->ici2, b->xfc, QX(3), b->qi0, b->qi1,
                       b->x1);
  }
  if (b->spectral_nudge) {                                     /* :497-504 */
    b->tspectral = b->tspectral + c->g->dtsec;
    if ((int)fmod(b->tspectral, b->dtrad) == 0) {
      /* the theta call sums over j = jci1..jci1 only, as the reference does */
      mospectral_nudge(c, b, c->jce1, c->jce2, c->ice1, c->ice2, c->jci1, c->jci1, c->ici1,
                       c->ici2, f->t, b->t0, b->t1);
      mospectral_nudge(c, b, c->jde1, c->jde2, c->ice1, c->ice2, c->jdi1, c->jdi2, c->ici1,
                       c->ici2, f->u, b->u0, b->u1);
      mospectral_nudge(c, b, c->jce1, c->jce2, c->ide1, c->ide2, c->jci1, c->jci2, c->idi1,
                       c->idi2, f->v, b->v0, b->v1);
    }
  }
  uvstagtouvx(c);
  temp_to_tvirt(c);
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(f->tetav, j, i, k) = A3(f->tvirt, j, i, k) / A3(f->pai, j, i, k);
}

static void set_threads(const orc_grid *g) {
#ifdef _OPENMP
  if (g->nthreads > 0) omp_set_num_threads(g->nthreads);
#else
  (void)g;
#endif
}

int orc_moloch_steps(const orc_grid *g, orc_fields *f, int nsteps) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  for (int n = 0; n < nsteps; ++n) {
    reset_tendencies(c);
    dynamical_core(c);
    diagnostics(c);
    extrapolate_surface_pressure(c);
    status_update(c, g->dtsec);
  }
  ctx_free(c);
  return 0;
}

int orc_boundary(const orc_grid *g, orc_fields *f, orc_bdy *b) {
  if (g->crm) return -1;            /* no lateral boundary on a periodic domain */
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  boundary(c, b);
  ctx_free(c);
  return 0;
}

int orc_moloch_steps_lbc(const orc_grid *g, orc_fields *f, orc_bdy *b,
                         const double *x1, int nsteps) {
  if (g->crm) return -1;
  ctx cc, *c = &cc;
  set_threads(g);
  /* with tracers boundary() also calls the chemistry module's
   * morelax_chiten (Main/mod_moloch.F90:495-497), not restated here */
  if (g->ntr > 0) return -1;
  if (ctx_init(c, g, f) != 0) return -1;
  for (int n = 0; n < nsteps; ++n) {
    reset_tendencies(c);
    dynamical_core(c);
    b->x1 = x1[n];
    boundary(c, b);
    diagnostics(c);
    extrapolate_surface_pressure(c);
    status_update(c, g->dtsec);
  }
  ctx_free(c);
  return 0;
}

/* mkslice, MOLOCH branch: Main/mod_slice.F90:112-160 and the tropopause /
 * PBL-top search :340-382 (icldmstrat, irceideal, ibltyp as given) */
int orc_mkslice(const orc_grid *g, orc_fields *f, orc_slice *sl) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  const int kz = c->kz, kzp1 = kz + 1, kzm1 = kz - 1;
  const double rovcp = c_rovcp;
  const double anorth[6] = {7.9925, 8.3329, 24.1731, -1.8069, 0.1082, -0.1493};
  const double asouth[6] = {8.1797, 8.1455, -23.4839, 1.1464, 0.0798, -0.1491};
  const double twopi = c_mathpi * 2.0;
#define SL3(a, j, i, k) A3(a, j, i, k)
  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j) A3(sl->pf, j, i, kzp1) = A2(f->ps, j, i);
  for (int k = 2; k <= kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(sl->pf, j, i, k) =
            c_p00 * pow(0.5 * (A3(f->pai, j, i, k) + A3(f->pai, j, i, k - 1)), c_cpovr);
  for (int i = c->ice1; i <= c->ice2; ++i)
    for (int j = c->jce1; j <= c->jce2; ++j)
      A3(sl->pf, j, i, 1) = A3(sl->pf, j, i, 2) - c_egrav * A3(f->rho, j, i, 1) *
                                                      (A3(sl->zetaf, j, i, 1) - A3(sl->zetaf, j, i, 2));
  for (int i = c->ici1; i <= c->ici2; ++i)
    for (int j = c->jci1; j <= c->jci2; ++j) {
      A2(sl->rhox2d, j, i) = A2(f->ps, j, i) / (c_rgas * A3(f->t, j, i, kz));
      A2(sl->tp2d, j, i) =
          A3(f->t, j, i, kz) * pow(A2(f->ps, j, i) / A3(f->p, j, i, kz), rovcp);
    }
  for (int k = 1; k <= kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j)
        A3(sl->th3d, j, i, k) = A3(f->t, j, i, k) * pow(c_p00 / A3(f->p, j, i, k), rovcp);
  for (int n = 1; n <= g->nqx; ++n) {
    double *q = QX(n);
    for (int k = 1; k <= kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j)
          if (A3(q, j, i, k) < qxcheckval[n - 1]) A3(q, j, i, k) = qxzeroval[n - 1];
  }
  for (int k = 1; k <= kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(sl->rhb3d, j, i, k) =
            fmind(fmaxd(A3(QX(1), j, i, k) / A3(f->qsat, j, i, k), sl->rhmin), sl->rhmax);
  for (int k = 1; k <= kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        A3(sl->wpx3d, j, i, k) = -c_egrav * A3(f->rho, j, i, k) * 0.5 *
                                 (A3(f->w, j, i, k + 1) + A3(f->w, j, i, k));
  if (sl->icldmstrat == 1)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        A2(sl->th700, j, i) = A3(sl->th3d, j, i, kz);
        for (int k = 2; k <= kzm1; ++k)
          if (A3(f->p, j, i, k) > 70000.0) {
            const double w1 = (A3(f->p, j, i, k) - 70000.0) /
                              (A3(f->p, j, i, k) - A3(f->p, j, i, k - 1));
            const double w2 = 1.0 - w1;
            A2(sl->th700, j, i) = A3(sl->th3d, j, i, k - 1) * w1 + A3(sl->th3d, j, i, k) * w2;
            break;
          }
      }
  if (sl->irceideal != 1)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j) {
        const double lat = A2(sl->xlat, j, i);
        const double *a = lat > 0.0 ? anorth : asouth;
        const double ztrop = a[0] + a[1] / pow(1.0 + exp(-(lat - a[2]) / a[3]), a[4]) +
                             a[5] * cos((twopi * (sl->calday - 28.0)) / sl->dayspy);
        A2(sl->ptrop, j, i) = c_p00 * exp(-ztrop / 8.4);
      }
  for (int i = 1; i <= c->iy; ++i)
    for (int j = 1; j <= c->jx; ++j) A2(sl->ktrop, j, i) = kz;
  for (int i = c->ici1; i <= c->ici2; ++i)
    for (int j = c->jci1; j <= c->jci2; ++j)
      for (int k = kzm1; k >= 2; --k) {
        A2(sl->ktrop, j, i) = k;
        if (A3(f->p, j, i, k) < A2(sl->ptrop, j, i)) break;
      }
  if (sl->ibltyp == 1) {
    for (int i = 1; i <= c->iy; ++i)
      for (int j = 1; j <= c->jx; ++j) A2(sl->kmxpbl, j, i) = kz;
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        for (int k = kzm1; k >= 2; --k) {
          if (A3(f->zeta, j, i, k) > 5000.0) break;
          A2(sl->kmxpbl, j, i) = k;
        }
  }
#undef SL3
  ctx_free(c);
  return 0;
}

int orc_dynamical_core(const orc_grid *g, orc_fields *f) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  dynamical_core(c);
  ctx_free(c);
  return 0;
}

int orc_status_update(const orc_grid *g, orc_fields *f, double dtinc) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  status_update(c, dtinc);
  ctx_free(c);
  return 0;
}

/* One sound() call (Main/mod_moloch.F90:545-736) with dts = dtsec/(nadv
 * nsound), for pinning its pieces.  stop = 0 runs the whole call; stop =
 * 1, 2, 3 return inside the first sub-step right after zdiv2 is formed
 * (:602-618), after divergence_damping and divergence_diffusion (:620-625),
 * and after the vertical-divergence term (:627-630).  zdiv2_out (kz
 * levels) and s_out (kz+1 levels), full extent, may be null. */
int orc_sound(const orc_grid *g, orc_fields *f, int stop, double *zdiv2_out,
              double *s_out) {
  ctx cc, *c = &cc;
  if (stop < 0 || stop > 3) return -1;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  c->probe_stop = stop;
  sound(c, c->dtsound);
  if (zdiv2_out) memcpy(zdiv2_out, c->zdiv2, P2(c) * (size_t)c->kz * sizeof(double));
  if (s_out) memcpy(s_out, c->s, P2(c) * (size_t)c->kzp1 * sizeof(double));
  ctx_free(c);
  return 0;
}

/* The curvature terms of advection alone (:811-825) on ux, vx. */
int orc_curvature(const orc_grid *g, orc_fields *f, double dta) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  curvature(c, dta);
  ctx_free(c);
  return 0;
}

/* The diagnostics of moloch() after the dynamical core: p, rho, qsat
 * (:348-352) and extrapolate_surface_pressure (:1592-1606). */
int orc_diagnostics(const orc_grid *g, orc_fields *f) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  diagnostics(c);
  extrapolate_surface_pressure(c);
  ctx_free(c);
  return 0;
}

/* The MOLOCH branch of init on a restart (Main/mod_init.F90:941-955): from
 * the SAV variables t, pai and qv (Main/mod_savefile.F90:618-628) it forms
 * p, qsat and rho as the diagnostics do, then tvirt = t*(1 + ep1*qv) -- water
 * vapour only, whatever ipptls -- and tetav = tvirt/pai on every cross point.
 * pf (:956-967) belongs to mkslice and is not held here. */
int orc_init_restart(const orc_grid *g, orc_fields *f) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  diagnostics(c);
  OMP_FOR
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ice1; i <= c->ice2; ++i)
      for (int j = c->jce1; j <= c->jce2; ++j) {
        size_t idx = &A3(f->t, j, i, k) - f->t;
        f->tvirt[idx] = f->t[idx] * (1.0 + c_ep1 * f->qx[idx]);
        f->tetav[idx] = f->tvirt[idx] / f->pai[idx];
      }
  ctx_free(c);
  return 0;
}

/* which = 0: temp_to_tvirt (:1608-1627); 1: tvirt_to_temp (:1629-1648). */
int orc_tv_convert(const orc_grid *g, orc_fields *f, int which) {
  ctx cc, *c = &cc;
  if (which != 0 && which != 1) return -1;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  if (which == 0)
    temp_to_tvirt(c);
  else
    tvirt_to_temp(c);
  ctx_free(c);
  return 0;
}

/* Zita levels and the gzita decay profile as init_moloch builds them
 * (Main/mod_params.F90:2460-2462, Main/mod_moloch.F90:273-274). */
int orc_zita_levels(int kz, double ztop, double a0, double *zitaf,
                    double *zitah, double *gzitaf, double *gzitah) {
  if (kz < 2) return -1;
  model_zitaf(zitaf, kz, ztop);
  model_zitah(zitah, kz, ztop);
  for (int k = 0; k <= kz; ++k) gzitaf[k] = gzita(zitaf[k], ztop, a0);
  for (int k = 0; k < kz; ++k) gzitah[k] = gzita(zitah[k], ztop, a0);
  return 0;
}

/* One wafone() call on field pp with a caller-supplied generalized vertical
 * velocity s (kz+1 levels, full extent). */
int orc_wafone(const orc_grid *g, orc_fields *f, const double *s, double *pp,
               double dta) {
  ctx cc, *c = &cc;
  set_threads(g);
  if (ctx_init(c, g, f) != 0) return -1;
  memcpy(c->s, s, P2(c) * (size_t)c->kzp1 * sizeof(double));
  wafone(c, pp, dta);
  ctx_free(c);
  return 0;
}

/* The column solve of :634-664 for one column of n = kz-1 unknowns, in the
 * reference's recurrences: forward from the top-most unknown (index n-1,
 * level kz) down to index 0 (level 2) with wwkw above the column zero and
 * w above the column = wtop; then back substitution upwards with w below
 * the column = 0.  zu/zd/rhs are indexed like wout (0 = level 2). */
int orc_thomas(int n, const double *zu, const double *zd, const double *rhs,
               double wtop, double *wout) {
  if (n < 1) return -1;
  column_solve(n, zu, zd, rhs, wtop, 0.0, 0.0, wout);
  return 0;
}

/* b(r) of the WAF limiter applied to local_flow_param(num, den). */
double orc_waf_b(double num, double den) {
  return waf_limiter(local_flow_param(num, den));
}

/* xkdamp and xknu profiles of init_moloch (Main/mod_moloch.F90:294-299). */
int orc_damping_profiles(int kz, double *xkdamp, double *xknu) {
  if (kz < 2) return -1;
  for (int k = 1; k <= kz; ++k) {
    xkdamp[k - 1] = 0.125 * 0.850 *
        (1.0 / ((double)k + 1.0) - 1.0 / ((double)kz + 2.0));
    xknu[k - 1] = 0.125 * (0.55 + 0.45 *
        ((double)(kz - k + 1) - 1.0) / ((double)kz - 1.0));
  }
  return 0;
}

/* Staggering operators: which = 0 uvstagtouvx (u,v -> ux,vx), 1
 * uvxtouvstag (ux,vx -> u,v), 2 zstagtoh (w -> tvirt as output), 3
 * htozstag (tvirt -> w). */
int orc_stagger(const orc_grid *g, orc_fields *f, int which) {
  ctx cc, *c = &cc;
  size_t p2 = (size_t)g->jx * (size_t)g->iy;
  double *w1 = malloc(p2 * sizeof(double));
  if (!w1) return -1;
  memcpy(w1, f->w, p2 * sizeof(double));
  if (ctx_init(c, g, f) != 0) { free(w1); return -1; }
  memcpy(f->w, w1, p2 * sizeof(double));   /* keep the caller's w(:,:,1) */
  free(w1);
  switch (which) {
    case 0: uvstagtouvx(c); break;
    case 1: uvxtouvstag(c); break;
    case 2: zstagtoh(c, f->w, f->tvirt); break;
    case 3: htozstag(c, f->tvirt, f->w); break;
    default: ctx_free(c); return -1;
  }
  ctx_free(c);
  return 0;
}

/* massck, MOLOCH branch: Main/mod_massck.F90:77-160 and :307-308 with the
 * reference's serial loop order (k outermost, j innermost) in FP64.  The
 * boundary-flux terms use the model time step dtsec (`dt` there). */
int orc_massck(const orc_grid *g, const orc_fields *f, orc_mass *m) {
  ctx cc, *c = &cc;
  if (ctx_init(c, g, (orc_fields *)f) != 0) return -1;
  const double dx = g->dx, dt = g->dtsec, dxsq = dx * dx;
  double tdrym = 0.0, tdadv = 0.0, tqmass = 0.0, tqadv = 0.0;
  for (int k = 1; k <= c->kz; ++k)
    for (int i = c->ici1; i <= c->ici2; ++i)
      for (int j = c->jci1; j <= c->jci2; ++j)
        tdrym = tdrym + dxsq * A3(f->dz, j, i, k) * A3(f->rho, j, i, k);
  /* boundary input: only where an edge is a physical boundary (has_bdy*) */
  if (!c->crm) {
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        tdadv = tdadv + A3(f->u, c->jde1, i, k) * dt * dx *
                            A3(f->dz, c->jce1, i, k) * A3(f->rho, c->jce1, i, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        tdadv = tdadv - A3(f->u, c->jde2, i, k) * dt * dx *
                            A3(f->dz, c->jce2, i, k) * A3(f->rho, c->jce2, i, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int j = c->jci1; j <= c->jci2; ++j)
        tdadv = tdadv + A3(f->v, j, c->ide1, k) * dt * dx *
                            A3(f->dz, j, c->ice1, k) * A3(f->rho, j, c->ice1, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int j = c->jci1; j <= c->jci2; ++j)
        tdadv = tdadv - A3(f->v, j, c->ide2, k) * dt * dx *
                            A3(f->dz, j, c->ice2, k) * A3(f->rho, j, c->ice2, k);
  }
  for (int n = 1; n <= g->nqx; ++n) {
    const double *q = QX(n);
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ici1; i <= c->ici2; ++i)
        for (int j = c->jci1; j <= c->jci2; ++j)
          tqmass = tqmass + A3(q, j, i, k) * dxsq * A3(f->dz, j, i, k) *
                                A3(f->rho, j, i, k);
    if (c->crm) continue;
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        tqadv = tqadv + A3(q, c->jce1, i, k) * A3(f->u, c->jde1, i, k) * dt *
                            dx * A3(f->dz, c->jce1, i, k) *
                            A3(f->rho, c->jce1, i, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int i = c->ice1; i <= c->ice2; ++i)
        tqadv = tqadv - A3(q, c->jce2, i, k) * A3(f->u, c->jde2, i, k) * dt *
                            dx * A3(f->dz, c->jce2, i, k) *
                            A3(f->rho, c->jce2, i, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int j = c->jci1; j <= c->jci2; ++j)
        tqadv = tqadv + A3(q, j, c->ice1, k) * A3(f->v, j, c->ide1, k) * dt *
                            dx * A3(f->dz, j, c->ice1, k) *
                            A3(f->rho, j, c->ice1, k);
    for (int k = 1; k <= c->kz; ++k)
      for (int j = c->jci1; j <= c->jci2; ++j)
        tqadv = tqadv - A3(q, j, c->ice2, k) * A3(f->v, j, c->ide2, k) * dt *
                            dx * A3(f->dz, j, c->ice2, k) *
                            A3(f->rho, j, c->ice2, k);
  }
  m->drymass = tdrym;
  m->qmass = tqmass;
  m->dryadv = tdadv;
  m->qadv = tqadv;
  ctx_free(c);
  return 0;
}
