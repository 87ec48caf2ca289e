/*
 * moloch_hip.h - C-ABI of libmoloch_hip.so, the MI355X (gfx950) MOLOCH
 * dynamical-core step.
 *
 * The reference seam is Main/mod_moloch.F90:127
 *     public :: allocate_moloch, init_moloch, moloch
 * three argument-less subroutines working on module state (mo_atm, mddom,
 * sfs).  This library keeps that state resident in HBM; a Fortran
 * ISO_C_BINDING module (regcm_amd/fortran/mod_moloch_gpu.F90) calls the
 * entry points below from the places the reference calls its subroutines.
 *
 * Conventions
 *  - Every entry returns 0 (MLC_OK) or a negative MLC_E_* status; the
 *    Fortran wrapper turns a non-zero status into fatal(), as the reference
 *    does (Main/abort.F90:17-33).  mlc_last_error() gives the message.
 *  - Host arrays are Fortran arrays with bounds lb(1:4)..ub(1:4) over
 *    (j, i, k, n): j contiguous, then i, then k, then species n.  2-D fields
 *    use k = n = 1; fields without species use n = 1.  The box must lie in
 *    the field's extent: j in 1..jx, i in 1..iy, k in 1..nk (nk = kz or kz+1).
 *  - One HIP stream per handle; calls come from one host thread
 *    (MPI_THREAD_FUNNELED in the reference, Main/regcm.F90:59).
 *  - Single domain with fixed (held) lateral boundaries: the mpplib halo
 *    exchanges of the reference are no-ops there
 *    (Main/mpplib/mod_mppparam.F90:1262-1316) and have no counterpart here.
 */
#ifndef MOLOCH_HIP_H
#define MOLOCH_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MLC_OK 0
#define MLC_E_ARG (-1)       /* bad argument or grid description */
#define MLC_E_HIP (-2)       /* HIP runtime error */
#define MLC_E_BOUNDS (-3)    /* host box outside the field's extent */
#define MLC_E_NONFINITE (-4) /* NaN/Inf found by mlc_check_finite */
#define MLC_E_STATE (-5)     /* call out of order (e.g. step before init) */

#define MLC_FP64 0   /* plain FP64: every stage in double, bit-reproducible */
#define MLC_MIXED 1  /* FP32 stencil bulk, FP64 column solve and Exner */

/* Field identifiers (state, tendencies, static metrics). */
enum mlc_field {
  MLC_U = 1,    /* mo_atm%u     (jde, ice, kz)            */
  MLC_V,        /* mo_atm%v     (jce, ide, kz)            */
  MLC_W,        /* mo_atm%w     (jce, ice, kz+1)          */
  MLC_PAI,      /* mo_atm%pai   (jce, ice, kz)            */
  MLC_TETAV,    /* mo_atm%tetav                           */
  MLC_T,        /* mo_atm%t                               */
  MLC_TVIRT,    /* mo_atm%tvirt                           */
  MLC_QX,       /* mo_atm%qx    (jce, ice, kz, nqx)       */
  MLC_UX,       /* mo_atm%ux                              */
  MLC_VX,       /* mo_atm%vx                              */
  MLC_P,        /* mo_atm%p                               */
  MLC_RHO,      /* mo_atm%rho                             */
  MLC_QSAT,     /* mo_atm%qs                              */
  MLC_PS,       /* sfs%psa      (jce, ice) 2-D            */
  MLC_TTEN,     /* mo_atm%tten  (jci, ici, kz)            */
  MLC_UTEN,
  MLC_VTEN,
  MLC_QXTEN,    /* (jci, ici, kz, nqx)                    */
  MLC_TKE = 30, /* mo_atm%tke   (jce, ice, kz+1), ibltyp 2 */
  MLC_TKETEN,   /* mo_atm%tketen (jci, ici, kz+1)         */
  MLC_TRAC,     /* mo_atm%trac  (jce, ice, kz, ntr)       */
  MLC_CHITEN,   /* mo_atm%chiten (jci, ici, kz, ntr)      */
  MLC_FMZ = 100,/* mo_atm%fmz   (jce, ice, kz)            */
  MLC_FMZF,     /* mo_atm%fmzf  (jce, ice, kz+1)          */
  MLC_RFMZU,    /* mo_atm%rfmzu (jde, ice, kz)            */
  MLC_RFMZV,    /* mo_atm%rfmzv (jce, ide, kz)            */
  MLC_BDYWTU,   /* bdywtu (jdi, ici, kz), setup_bdywt     */
  MLC_BDYWTV,   /* bdywtv (jci, idi, kz)                  */
  MLC_BDYWTW,   /* bdywtw (jci, ici, kz)                  */
  MLC_ZETA,     /* mo_atm%zeta  (jce, ice, kz)            */
  MLC_DZ,       /* mo_atm%dz    (jce, ice, kz)            */
  MLC_MX,       /* mddom%msfx   (jde, ide) 2-D            */
  MLC_MU,       /* mddom%msfu                             */
  MLC_MV,       /* mddom%msfv                             */
  MLC_HX,       /* mddom%hx     (jde, ice)                */
  MLC_HY,       /* mddom%hy     (jce, ide)                */
  MLC_CORU,     /* coru         (jde, ice)                */
  MLC_CORV,     /* corv         (jce, ide)                */
  MLC_RLAT,     /* mddom%rlat   (ide) 1-D, as (1, iy)     */
  MLC_FFILT     /* ffilt        (kz) 1-D, as (1, 1, kz)   */
};

/* Surface geopotential (m2/s2) that compute_moloch_static forms the
 * terrain metrics from (Main/mod_params.F90:3356-3385); optional statics,
 * see MLC_PARAM_MO_H. */
enum mlc_terrain_field {
  MLC_HT = 122,   /* mddom%ht  (jce, ice) 2-D */
  MLC_HTU,        /* mddom%htu (jde, ice)     */
  MLC_HTV         /* mddom%htv (jce, ide)     */
};

typedef struct mlc_grid_desc {
  int jx, iy, kz;      /* dot-point extents (Share/mod_dynparam.F90) and levels */
  int nqx;             /* species in qx (Main/mod_params.F90:1724-1737) */
  int ipptls;          /* 0, 1 or 2: species entering Tv and advection */
  int nadv, nsound;    /* mo_nadv, mo_nsound */
  int do_divdamp;      /* mo_divdamp */
  int do_divfilter;    /* mo_divfilter */
  int lrotllr;         /* iproj == 'ROTLLR' */
  int precision;       /* MLC_FP64 or MLC_MIXED */
  int ibltyp;          /* PBL scheme; 2 (UW) carries TKE through the step */
  int ntr;             /* chemical tracers (ichem == 1 when > 0), 0 .. 32 */
  int crm;             /* i_crm = 1 (which implies i_band): doubly periodic,
                          jx x iy cross points, no lateral boundary
                          (Main/mpplib/mod_mppparam.F90:1290-1309) */
  int reserved[2];     /* must be zero */
  double dx;           /* grid spacing (m) */
  double dtsec;        /* model time step (s) */
  double mo_ztop;      /* model top (m) */
  double mo_a0;        /* gzita shape parameter */
} mlc_grid_desc;

typedef struct mlc_handle mlc_handle;

/* allocate_moloch (Main/mod_moloch.F90:159-199): device state for one
 * domain, zero-filled like getmem (Share/mod_space.F90:309-310). */
int mlc_create(const mlc_grid_desc *g, mlc_handle **h);
int mlc_destroy(mlc_handle *h);

/* Upload a static metric (mo_atm%fmz ..., mddom%msfx ..., ffilt); replaces
 * the assignpnt aliases of init_moloch (Main/mod_moloch.F90:204-255). */
int mlc_set_static(mlc_handle *h, int field, const double *host,
                   const int lb[4], const int ub[4]);
/* Upload / download a state or tendency field (mo_atm%u ...). */
int mlc_put_state(mlc_handle *h, int field, const double *host,
                  const int lb[4], const int ub[4]);
int mlc_get_state(mlc_handle *h, int field, double *host, const int lb[4],
                  const int ub[4]);

/* init_moloch (Main/mod_moloch.F90:201-308): derived metrics mx2, rmx,
 * rmu, rmv, gzitak, gzitakh, xkdamp, xknu, rdzita, dtstepa, dtsound, and
 * w(:,:,1) = 0.  Call after every static field has been set. */
int mlc_init(mlc_handle *h);

/* moloch() (Main/mod_moloch.F90:312-446) with physics off and fixed
 * lateral boundaries, repeated nsteps times: reset_tendencies,
 * dynamical_core, p/rho/qsat, extrapolate_surface_pressure,
 * status_update(dtsec).  Asynchronous on the handle's stream. */
int mlc_step(mlc_handle *h, int nsteps);

/* The pieces of moloch() for a caller that runs physics on the host
 * between them: reset_tendencies (:1044-1083), dynamical_core
 * (:1085-1141), the diagnostics of :348-354 with
 * extrapolate_surface_pressure (:1592-1606), status_update (:1403-1443). */
int mlc_reset_tendencies(mlc_handle *h);
int mlc_dynamical_core(mlc_handle *h);
int mlc_diagnostics(mlc_handle *h);
int mlc_status_update(mlc_handle *h, double dtinc);

/* Run-time scalars of the step.  MLC_PARAM_TKEMIN: the PBL scheme's
 * tkemin (Main/mod_pbl_interface.F90:50, set by the scheme's init), the
 * floor of status_update (Main/mod_moloch.F90:1421-1422) and the inflow
 * value of the TKE boundary rows (Main/mod_bdycod.F90:1684-1694); 0 until
 * set. */
#define MLC_PARAM_TKEMIN 1
/* MLC_PARAM_MO_H: mo_h (m), the decay height of the zita coordinate's
 * terrain term (zfz, bzitap, Share/mod_zita.F90:88-93, 102-107).  Optional: with it and
 * the surface geopotential (MLC_HT, MLC_HTU, MLC_HTV), mlc_init forms fmz,
 * fmzf, rfmzu and rfmzv as compute_moloch_static does (md_fmz,
 * Main/mod_params.F90:3356-3385) and compares them with the uploaded arrays
 * bit for bit; where every value agrees, the kernels form the metrics from
 * the 2-D orography and two per-level tables instead of streaming the four
 * 3-D arrays (MLC_VAR_TERRAIN).  Call mlc_init again after setting it. */
#define MLC_PARAM_MO_H 2
int mlc_set_param(mlc_handle *h, int param, double value);

/* massck, MOLOCH branch (Main/mod_massck.F90:77-160, 307-308): out[0] dry
 * air mass, out[1] water mass, out[2] dry-air boundary flux, out[3] water
 * boundary flux, in kg, as compensated FP64 sums.  Synchronous. */
int mlc_massck(mlc_handle *h, double out[4]);

/* mlc_massck without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory at the point of the call
 * (tight loops, harnesses that forbid hipMalloc).  The same block partials
 * and the same combine order as mlc_massck, so the same bits; the partials
 * and totals live in a static device array of the library (8 MiB per GPU,
 * held until the process ends), not in the handle's reduction buffer.
 * Domains whose budget needs more than 131072 blocks of partials (64 x 4
 * columns per block, one level each: e.g. more than 1024 x 512 x 64 points)
 * are refused with MLC_E_ARG.  The current device must be the one the
 * handle was created on.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (several host threads, or handles on
 * different streams) race.  Synchronous. */
int mlc_massck_noalloc(mlc_handle *h, double out[4]);

/* CFL/NaN guard of Main/mod_moloch.F90:407-422: max/min of ps over the
 * interior; returns MLC_E_NONFINITE when either is NaN or Inf.
 * Synchronous. */
int mlc_check_finite(mlc_handle *h, double *psmax, double *psmin);

/* Wait for all work queued on the handle's stream. */
int mlc_synchronize(mlc_handle *h);

/* Per-kernel timing and reset of the totals: enable = 0 off, 1 every kernel
 * slot, or a mask with bit (slot + 1) set for each slot whose launches alone
 * are timed (2 << slot for one slot).  A timed launch records a start and a
 * stop HIP event itself (hipExtLaunchKernel), so the time is the kernel's,
 * without the dispatch gap; results are unchanged.  mlc_kernel_stats gives
 * the accumulated launches and milliseconds of kernel slot `slot`
 * (0..mlc_kernel_count()-1).  No counterpart in the reference (its timers
 * are host-side, Main/mpplib/mod_timer.F90). */
int mlc_profile(mlc_handle *h, int enable);
int mlc_kernel_count(void);
const char *mlc_kernel_name(int slot);
int mlc_kernel_stats(mlc_handle *h, int slot, long long *launches,
                     double *total_ms);

/* Launch-shape variants of a stage.  Every variant of a stage gives
 * bit-identical results; their speed depends on the domain shape.  No
 * reference counterpart (a GPU launch choice). */
#define MLC_VAR_WAF_H 0       /* horizontal WAF: 0 row-streaming, 1 LDS tile (default) */
#define MLC_VAR_WAF_ROWS 1    /* row-streaming WAF: i rows per wavefront, 4..256 (16) */
#define MLC_VAR_WAF_GROUP 2   /* row-streaming WAF: fields per lane, 2 or 4 (4) */
#define MLC_VAR_LEVEL_CHUNK 3 /* divergence (and momentum, unless MLC_VAR_MOM_CHUNK is set):
                                 levels per thread, 1..kz (8) */
#define MLC_VAR_COLUMN 4      /* column solve: 0 LDS three-phase, 1 streamed, 2 LDS with
                                 level chunks in registers (default; kz > 64 runs 0) */
#define MLC_VAR_WAFV_GROUP 5  /* vertical WAF: fields per thread, 1, 2 or 4 (4) */
#define MLC_VAR_XCD_MAP 6     /* single-domain launches: 1 XCD-contiguous tile order (default), 0 plain */
#define MLC_VAR_WAF_TILE 7    /* LDS-tile WAF: output rows per tile, 4 or 8 (default 8 in FP64, 4 mixed) */
#define MLC_VAR_WAF_HGROUP 8  /* LDS-tile WAF: fields per block, 0 = the whole batch (default) */
#define MLC_VAR_COLUMN_J 9    /* chunked column solve: columns per block, 16, 32 (default) or
                                 64 (up to 42 levels) */
#define MLC_VAR_MIXED_MERGE 10 /* mixed form: float and double WAF batches in one launch,
                                 bit 0 vertical (default 1), bit 1 horizontal as two block
                                 ranges, bit 2 horizontal in the same blocks (one coefficient
                                 setup; whole batch per block) */
#define MLC_VAR_TERRAIN 11    /* terrain metrics formed from the orography instead of streamed
                                 from the uploaded 3-D arrays, as a mask of stages: 1
                                 divergence, 2 column solve, 4 advection staggering, 8
                                 horizontal WAF.  mlc_init sets 7 (FP64) or 6 (mixed) when the
                                 formed values equal every uploaded one (MLC_PARAM_MO_H), else
                                 0; a non-zero mask otherwise is MLC_E_STATE.  Same bits
                                 either way. */
#define MLC_VAR_MOM_CHUNK 12  /* momentum update: levels per thread, 1..kz, or 0 for
                                 MLC_VAR_LEVEL_CHUNK's (default kz: whole columns) */
#define MLC_VAR_STAGGER_CHUNK 13 /* staggering passes (winds to and from the advection
                                 grid): levels per thread, 1..kz (1) */
#define MLC_VAR_COUNT 14
int mlc_set_variant(mlc_handle *h, int stage, int value);
int mlc_get_variant(const mlc_handle *h, int stage);

/* Autotuner: times whole steps of this domain under the launch-shape
 * variants (coordinate descent over the stages above, each candidate timed
 * over nsteps steps after one warm-up step) and keeps a candidate only when
 * it is at least 1% faster than the value it would replace.  The
 * state is snapshotted on the device first and restored afterwards, so the
 * call leaves the domain exactly as it found it.  *ms_per_step (may be
 * null) receives the best step time.  Synchronous. */
int mlc_autotune(mlc_handle *h, int nsteps, double *ms_per_step);

/* Number of grid points of the step metric: the cross points times kz,
 * (jx-1)(iy-1)kz, or jx*iy*kz on a doubly periodic domain. */
long long mlc_npoints(const mlc_handle *h);

/* The value of a physical constant as the kernels use it (NaN if unknown):
 * for checking against Share/mod_constants.F90. */
double mlc_constant(const char *name);

/* ---- Real-case lateral boundaries --------------------------------------
 * boundary() of Main/mod_moloch.F90:448-529, run when do_apply_bdy
 * (:305): bdyval's MOLOCH branch (Main/mod_bdycod.F90:1638-1866; time
 * interpolation between the boundary states b0, b1 with x0 = 1 - x1,
 * inflow/outflow switch for the other water species and for w), optional
 * top nudging motopnudge (:4049-4081), the Davies relaxation morelax
 * (:3962-4031) on u, v, t, pai, qv, w, qc, qi, the spectral nudging of t,
 * u, v (mospectral_nudge :3940-3960), then uvstagtouvx, temp_to_tvirt and
 * tetav = tvirt/pai.  No tke, no tracers.
 *
 * After mlc_lbc_configure, upload the boundary states with mlc_put_state
 * (ids MLC_BND_*; full extents like the matching state field, 2-D for ps),
 * the Davies weights (MLC_RELAX_U/V/C: hefc(ibnd(j,i),k) of ba_ud, ba_vd,
 * ba_cr where ibnd > 0, -1 elsewhere) and, with top nudging, MLC_TNUDGE
 * (kz values of tnudge, k = 1..kz). */
typedef struct mlc_lbc_desc {
  int present_qc;      /* is_present_qc() */
  int present_qi;      /* is_present_qi() */
  int top_nudge;       /* mo_top_nudge */
  int nztop;           /* levels 1..nztop are nudged at the top */
  int spectral_nudge;  /* mo_spectral_nudge (bases from mlc_lbc_spectral) */
  int km, lm;          /* spectral nudging: 2km x-modes, 2lm y-modes */
  int reserved;        /* must be zero */
  double dtrad;        /* spectral nudging interval (s) */
} mlc_lbc_desc;

enum mlc_lbc_field {
  MLC_BND_PS0 = 200, MLC_BND_PS1, MLC_BND_U0, MLC_BND_U1, MLC_BND_V0, MLC_BND_V1,
  MLC_BND_T0, MLC_BND_T1, MLC_BND_PAI0, MLC_BND_PAI1, MLC_BND_QV0, MLC_BND_QV1,
  MLC_BND_QC0, MLC_BND_QC1, MLC_BND_QI0, MLC_BND_QI1,
  MLC_RELAX_U = 216, MLC_RELAX_V = 217, MLC_RELAX_C = 218, MLC_TNUDGE = 219
};

int mlc_lbc_configure(mlc_handle *h, const mlc_lbc_desc *desc);
/* Spectral-nudging bases of lowpass_init (Main/mod_bdycod.F90:3844-3896):
 * bvx(jx, 2km) and bvy(iy, 2lm) in Fortran order, cnudge(kz).  Nudging
 * runs inside the boundary stage whenever the accumulated step time is a
 * multiple of dtrad (Main/mod_moloch.F90:497-504). */
int mlc_lbc_spectral(mlc_handle *h, const double *bvx, const double *bvy,
                     const double *cnudge);
/* boundary() now, with time weight x1 (for a caller running the split
 * sequence: after mlc_dynamical_core, before mlc_diagnostics).  With
 * tracers (ntr > 0) the chemistry module's morelax_chiten
 * (Main/mod_che_interface.F90, called at Main/mod_moloch.F90:495-497) is
 * the caller's, between this call and mlc_status_update. */
int mlc_boundary(mlc_handle *h, double x1);
/* moloch() with do_apply_bdy, nsteps times; step n uses x1[n].  Refused
 * (MLC_E_ARG) with tracers, whose boundary relaxation is the caller's. */
int mlc_step_lbc(mlc_handle *h, int nsteps, const double *x1);

/* ---- Physics input slices ----------------------------------------------
 * mkslice, MOLOCH branch (Main/mod_slice.F90:112-160, 340-382), on the
 * device: pf (kz+1 full-level pressure), th3d, rhb3d, wpx3d (omega), rhox2d,
 * tp2d, th700 (icldmstrat = 1), ptrop (irceideal /= 1; 10000 Pa otherwise),
 * ktrop and kmxpbl (ibltyp = 1) as level numbers, and the qx clamp in
 * place.  atms%ps2d is the handle's ps.  Needs the statics MLC_ZETAF
 * (mo_atm%zetaf, kz+1) and, with irceideal /= 1, MLC_XLAT.  Outputs are read
 * with mlc_get_state (MLC_SL_*), always double. */
typedef struct mlc_slice_desc {
  double rhmin, rhmax;   /* namelist RH limits (Main/mod_params.F90:381-382) */
  double calday, dayspy; /* day of year and days per year of the calendar */
  int icldmstrat, irceideal, ibltyp, reserved;
} mlc_slice_desc;

enum mlc_slice_field {
  MLC_ZETAF = 120, MLC_XLAT = 121,
  MLC_SL_PF = 300, MLC_SL_TH = 301, MLC_SL_RH = 302, MLC_SL_OMEGA = 303,
  MLC_SL_RHOX = 304, MLC_SL_TP = 305, MLC_SL_TH700 = 306, MLC_SL_PTROP = 307,
  MLC_SL_KTROP = 308, MLC_SL_KMXPBL = 309
};

int mlc_mkslice(mlc_handle *h, const mlc_slice_desc *desc);

/* ---- Horizontal diffusion of the MM cores --------------------------------
 * Main/mod_diffusion.F90 as a device operator on its own handle (the MOLOCH
 * step does not call it; its own horizontal diffusion is the divergence
 * filter inside mlc_step).  One domain owning all four lateral boundaries;
 * FP64.  Fields are moved with mlc_diff_put / mlc_diff_get over the same
 * (j, i, k, n) boxes as mlc_put_state. */
typedef struct mlc_diff_desc {
  int jx, iy, kz;        /* dot-point extents and levels */
  int idiffu;            /* 1 del4 + del2 ring (default), 2 nine-point, 3 flux-limited 6th order */
  int idynamic;          /* 1 hydrostatic, 2 non-hydrostatic (coefficient uses w) */
  int diffu_hgtf;        /* 1: topographic reduction of the background coefficient */
  int nspec;             /* species capacity of MLC_DF_F / MLC_DF_FTEN */
  int reserved;          /* must be zero */
  double dx, dtsec;      /* grid spacing (m), time step (s) */
  double ckh, adyndif;   /* namelist dynparam (Main/mod_params.F90:803-804) */
} mlc_diff_desc;

typedef struct mlc_diff mlc_diff;

enum mlc_diff_field {
  MLC_DF_HT = 400,   /* mddom%ht (jx, iy) */
  MLC_DF_MSFD,       /* mddom%msfd (jx, iy) */
  MLC_DF_PC,         /* sfs%psb (jx, iy) */
  MLC_DF_PD,         /* sfs%psdotb (jx, iy) */
  MLC_DF_UD,         /* atms%ubd3d (jx, iy, kz): coefficient input and diffu_d's u */
  MLC_DF_VD,         /* atms%vbd3d */
  MLC_DF_WX,         /* atms%wb3d (jx, iy, kz+1) */
  MLC_DF_UTEN,       /* diffu_d output tendencies (jx, iy, kz) */
  MLC_DF_VTEN,
  MLC_DF_F,          /* diffu_x field: nspec species of kz+1 levels */
  MLC_DF_FTEN,       /* diffu_x tendency, same shape */
  MLC_DF_HGFACT,     /* hgfact (jx, iy), read back */
  MLC_DF_XKC,        /* xkc (jx, iy, kz), read back */
  MLC_DF_XKD,        /* xkd (jx, iy, kz), read back */
  MLC_DF_XKCF        /* xkcf (jx, iy, kz+1), read back */
};

/* diffu_x forms: diffu_x3d(ften, f) over kz levels; diffu_x3df(ften, f, fac)
 * over kz+1 levels; diffu_x4d(ften, f, n1, n2, fac) over kz levels. */
enum { MLC_DIFFU_X3D = 0, MLC_DIFFU_X3DF = 1, MLC_DIFFU_X4D = 2 };

/* allocate_mod_diffusion (Main/mod_diffusion.F90:90-97) and the scalars of
 * initialize_diffusion (:110-125, :152-153) */
int mlc_diff_create(const mlc_diff_desc *desc, mlc_diff **out);
int mlc_diff_destroy(mlc_diff *d);
int mlc_diff_put(mlc_diff *d, int field, const double *host, const int lb[4], const int ub[4]);
int mlc_diff_get(mlc_diff *d, int field, double *host, const int lb[4], const int ub[4]);
/* initialize_diffusion's hgfact (:131-146); needs MLC_DF_HT with diffu_hgtf = 1 */
int mlc_diff_initialize(mlc_diff *d);
/* calc_coeff (:166-248): xkc, xkd, xkcf from UD, VD, WX, PC, PD */
int mlc_diff_calc_coeff(mlc_diff *d);
/* diffu_d (:263-511): UTEN, VTEN += the diffusion of UD / msfd, VD / msfd */
int mlc_diff_d(mlc_diff *d);
/* diffu_x (:513-913) on species n1..n2 of F into FTEN; forms 0 and 1 treat
 * each species as one 3-D field */
int mlc_diff_x(mlc_diff *d, int form, int n1, int n2, double fac);
int mlc_diff_synchronize(mlc_diff *d);

/* ---- Restart ------------------------------------------------------------
 * The MOLOCH branch of init on a restart (Main/mod_init.F90:941-955).  A
 * restart run reads the SAV variables u, v, w, t, pai, qx, ps (and tke with
 * ibltyp 2) through RegCM's own read_savefile (Main/mod_savefile.F90:618-628),
 * the caller puts them with mlc_put_state and calls mlc_init; this entry then
 * forms p = pai^(cp/R)*p00, qsat, rho = p/(R*t), tvirt = t*(1 + ep1*qv)
 * (water vapour only, whatever ipptls) and tetav = tvirt/pai on every cross
 * point.  A save is mlc_get_state of the same variables into write_savefile.
 * Asynchronous on the handle's stream. */
int mlc_init_restart(mlc_handle *h);

/* Message of the last failed call on this thread. */
const char *mlc_last_error(void);

/* ---- Batched form ------------------------------------------------------
 * n independent domains (ensemble members, or domains of uneven jx, iy, kz,
 * dx, dtsec) advanced together: every stage of a step runs as ONE launch
 * over the tiles of all members, and the whole step is one hipGraph launch.
 * There is no reference counterpart (RegCM runs one domain per process);
 * each member behaves exactly as a domain of its own (mlc_create): the
 * per-member results are bit-identical to stepping it alone.  Members must
 * share precision, nadv and nsound.  They are reached through
 * mlc_batch_member() for mlc_set_static / mlc_put_state / mlc_get_state /
 * mlc_init / mlc_massck / mlc_check_finite; mlc_destroy refuses them. */
typedef struct mlc_batch mlc_batch;
int mlc_batch_create(const mlc_grid_desc *descs, int n, mlc_batch **b);
int mlc_batch_member(mlc_batch *b, int m, mlc_handle **h);
/* moloch() on every member, nsteps times: one graph launch per member
 * group and step (one per step with a single group). */
int mlc_batch_step(mlc_batch *b, int nsteps);
/* Split the members into k groups, 1 <= k <= 4 (the hardware queues of a
 * process), balanced by grid points, each with its own stream and step
 * graphs, so that the groups' stages overlap on the device; k = 1 (the
 * default) is one graph for the whole batch. */
int mlc_batch_set_streams(mlc_batch *b, int k);
/* mlc_autotune on every member (each member keeps its own launch shapes;
 * members with different kernels run as separate launches of a stage),
 * then the group count of mlc_batch_set_streams (1, 2, 4, at most the
 * member count) timed over nsteps steps; the state is left as it was. */
int mlc_batch_autotune(mlc_batch *b, int nsteps);
/* The current number of member groups (streams). */
int mlc_batch_streams(const mlc_batch *b);
int mlc_batch_synchronize(mlc_batch *b);
/* Total grid points of the step metric over all members. */
long long mlc_batch_npoints(const mlc_batch *b);
int mlc_batch_destroy(mlc_batch *b);

#ifdef __cplusplus
}
#endif
#endif
