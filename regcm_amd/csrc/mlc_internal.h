// Internal declarations of libmoloch_hip: device layout, the handle, the
// physical constants and the kernel launchers.
//
// Device layout (HBM): every 3-D field is a dense [nk][iy][pj] array of the
// storage type, j contiguous, with pj = jx rounded up to a multiple of 32 so
// that each (i,k) row starts on a 128/256-byte boundary.  Index (j,i,k) is
// 1-based as in the reference:  ((k-1)*iy + (i-1))*pj + (j-1).  Cross fields
// leave j = jx / i = iy unused.  2-D fields are [iy][pj].
//
// Forms: in MLC_FP64 every array is double.  In MLC_MIXED the "bulk" arrays
// (winds, water species, divergence, s, metrics, diagnostics) are float
// and the stencils on them compute in float; the Exner function pai, theta_v
// (tetav), w, tetavf, the column solve, the Exner update, the momentum
// arithmetic, the advection of pai and tetav, p and the budget sums stay
// double.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/moloch_hip.h"

namespace mlc {

// ---- constants: Share/mod_constants.F90 (non-RCEMIP), same expressions ----
constexpr double kEgrav = 9.80665;
constexpr double kRegrav = 1.0 / kEgrav;
constexpr double kRgasmol = 6.02214076e23 * 1.3806490e-23;
constexpr double kAmd = 28.96454;
constexpr double kAmw = 18.01528;
constexpr double kRgas = (kRgasmol / kAmd) * 1000.0;
constexpr double kCpd = 3.5 * kRgas;
constexpr double kCvd = 2.5 * kRgas;
constexpr double kRdrcv = kRgas / kCvd;
constexpr double kCpovr = kCpd / kRgas;
constexpr double kGovr = kEgrav / kRgas;
constexpr double kGovcp = kEgrav / kCpd;
constexpr double kEp1 = kAmd / kAmw - 1.0;
constexpr double kEp2 = kAmw / kAmd;
constexpr double kTzero = 273.15;
constexpr double kP00 = 1.0e5;
constexpr double kLrate = 0.00649;
constexpr double kMathpi = 3.14159265358979323846;
constexpr double kDegrad = kMathpi / 180.0;
constexpr double kEarthrad = 6.371229e6;
constexpr double kRearthrad = 1.0 / kEarthrad;
constexpr double kEomeg2 = 2.0 * 7.2921159e-5;

constexpr int kMaxAdv = 16;   // fields per WAF launch (its pointer tables)
constexpr int kMaxTr = 32;    // chemical tracers
// Advected-field slots: tetav, pai, ux, vx, wx, qx(1:nqx), then tkex
// (ibltyp 2) and trac(1:ntr).  Not every slot is advected (qx(2:nqx) are
// not with ipptls 0); the handle lists the advected ones.
enum AdvSlot { A_TETAV = 0, A_PAI = 1, A_UX = 2, A_VX = 3, A_WX = 4, A_QX0 = 5 };
constexpr int kMaxSlots = A_QX0 + 10 + 1 + kMaxTr;

// Geometry passed by value to every kernel.
struct Dims {
  int jx, iy, kz, pj;
  long long plane;   // iy * pj
  // Extents (Main/mod_atm_interface.F90:231-287).  Cross points j = 1..je2,
  // i = 1..ie2, their interior jc1..jc2 x ic1..ic2; the dot-point interior
  // is jc1..je2 (jdi) and ic1..ie2 (idi).  With lateral boundaries je2 =
  // jx-1, jc = 2..jx-2.  crm = 1: a doubly periodic domain (i_crm), every
  // extent is the whole domain (je2 = jc2 = jx, jc1 = 1) and a neighbour
  // index past an edge wraps to its periodic partner, which is what the
  // reference's exchanges put in the halos (Main/mpplib/mod_mppparam.F90:
  // 1298-1309).
  int crm;
  int je2, ie2, jc1, jc2, ic1, ic2;
  __host__ __device__ long long at(int j, int i, int k) const {
    return ((long long)(k - 1) * iy + (i - 1)) * (long long)pj + (j - 1);
  }
  __host__ __device__ long long at2(int j, int i) const {
    return (long long)(i - 1) * pj + (j - 1);
  }
  // periodic partner of an index within one period of the domain
  __host__ __device__ int wj(int j) const {
    return crm ? (j < 1 ? j + jx : (j > jx ? j - jx : j)) : j;
  }
  __host__ __device__ int wi(int i) const {
    return crm ? (i < 1 ? i + iy : (i > iy ? i - iy : i)) : i;
  }
  __host__ __device__ long long atw(int j, int i, int k) const { return at(wj(j), wi(i), k); }
  __host__ __device__ long long at2w(int j, int i) const { return at2(wj(j), wi(i)); }
  // element offsets of (j+s, i) from (j, i) and of (j, i+s) from (j, i)
  __host__ __device__ long long oj(int j, int s) const { return (long long)(wj(j + s) - j); }
  __host__ __device__ long long oi(int i, int s) const {
    return (long long)(wi(i + s) - i) * (long long)pj;
  }
};

// Dims of a domain of jx x iy dot points and kz levels.
inline Dims make_dims(int jx, int iy, int kz, int crm) {
  Dims d{};
  d.jx = jx;
  d.iy = iy;
  d.kz = kz;
  d.pj = (jx + 31) / 32 * 32;
  d.plane = (long long)iy * d.pj;
  d.crm = crm ? 1 : 0;
  d.je2 = crm ? jx : jx - 1;
  d.ie2 = crm ? iy : iy - 1;
  d.jc1 = crm ? 1 : 2;
  d.jc2 = crm ? jx : jx - 2;
  d.ic1 = crm ? 1 : 2;
  d.ic2 = crm ? iy : iy - 2;
  return d;
}

// A device array of 4- or 8-byte reals.
struct Buf {
  void *p = nullptr;
  int es = 8;
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// The lateral-boundary weights bdywt{u,v,w}(j,i,k) of setup_bdywt
// (Main/mod_bdycod.F90:4033-4047) are 1 - hefc(ibnd(j,i),k) or 1: every
// column (j,i) is one of a few distinct columns.  The kernels read them as
// tab[idx(j,i)][k], a 2-D index into a table of the distinct columns (exact
// copies of the uploaded values), instead of streaming a 3-D array.
template <typename R>
struct BdyWt {
  const int *idx;   // [iy][pj] column ids
  const R *tab;     // [ncol][kz]
  int kz;
  __host__ __device__ R at(long long c2, int k) const {
    return tab[(long long)idx[c2] * kz + (k - 1)];
  }
  // the column of (j,i): col(c2)[k - 1] == at(c2, k)
  __host__ __device__ const R *col(long long c2) const {
    return tab + (long long)idx[c2] * kz;
  }
};

// The terrain-following metrics of compute_moloch_static
// (Main/mod_params.F90:3356-3385) formed from the surface geopotential:
//   fmz(j,i,k)   = 1/(gzitap(zitah(k))*orog + bzitap(zitah(k))),  orog = ht*regrav
//   fmzf(j,i,k)  = the same on zita(k), k = 1..kz+1
//   rfmzu(j,i,k) = 1/md_fmz(zitah(k), htu(j,i)),  rfmzv on htv
// (md_fmz / md_fmz_h, Share/mod_zita.F90:135-141, 165-170), with the same
// operations in the same order as the reference, so each value is the one
// the caller computed.  gzitap and bzitap are per-level tables; ht, htu, htv
// are 2-D.  `on` is set only after mlc_init has compared the formula with
// every uploaded metric bit for bit, and then per stage; otherwise the
// kernels stream the uploaded 3-D arrays.
struct ZMet {
  const double *ga, *ba;   // gzitap, bzitap: zitah(1..kz) then zita(1..kz+1)
  const double *hc, *hu, *hv;   // ht, htu, htv (m2/s2), [iy][pj]
  int kz, on;
  __host__ __device__ double den(const double *h, long long c2, int t) const {
    return ga[t] * (h[c2] * kRegrav) + ba[t];
  }
  __host__ __device__ double fmz(long long c2, int k) const { return 1.0 / den(hc, c2, k - 1); }
  __host__ __device__ double fmzf(long long c2, int k) const {
    return 1.0 / den(hc, c2, kz + k - 1);
  }
  __host__ __device__ double rfmzu(long long c2, int k) const {
    return 1.0 / (1.0 / den(hu, c2, k - 1));
  }
  __host__ __device__ double rfmzv(long long c2, int k) const {
    return 1.0 / (1.0 / den(hv, c2, k - 1));
  }
  // rfmzu (h = hu) or rfmzv (h = hv)
  __host__ __device__ double rfm(const double *h, long long c2, int k) const {
    return 1.0 / (1.0 / den(h, c2, k - 1));
  }
};

// Kernel slots for per-kernel timing.
enum KernelSlot {
  K_DIV = 0, K_COL, K_MOM, K_STAG_IN, K_WAF_V, K_WAF_H, K_CURV, K_STAG_OUT,
  K_PS, K_TAIL, K_UVSTAG, K_FILL, K_FRAME, K_MASSCK, K_FINITE, K_LBC, K_SLICE, K_COUNT
};
extern const char *const kKernelNames[K_COUNT];

// One recorded launch of a stage: the batched entry point of its argument
// struct, its grid over one domain and the struct's bytes.
struct LaunchRec {
  int slot = 0;
  const void *many = nullptr;
  dim3 grid, block;
  size_t lds = 0;
  std::vector<unsigned char> args;
};

}  // namespace mlc

struct mlc_handle {
  mlc_grid_desc g{};
  mlc::Dims d{};
  hipStream_t stream = nullptr;
  bool mixed = false;            // MLC_MIXED form
  int es = 8;                    // element size of the bulk arrays
  int nadvf = 0;                 // number of advected fields F
  int adv_list[mlc::kMaxSlots] = {};   // their slots, in the reference's order
  int nslots = 0;                // slots in use
  int s_tkex = -1, s_tr0 = -1;   // slot of tkex (ibltyp 2) and of trac(1) (ntr > 0)
  double tkemin = 0.0;           // MLC_PARAM_TKEMIN
  bool inited = false;
  bool tend_dirty = false;       // tendencies may be non-zero
  unsigned config_gen = 0;       // bumped when a parameter or launch shape changes
  double rdx = 0, rdzita = 0, dtstepa = 0, dtsound = 0;

  // state (bulk type)
  mlc::Buf u, v, t, tvirt, p, rho, qsat, ps, tten, uten, vten, qxten;
  double *w = nullptr;           // always double
  // advected fields, double-buffered: adv[b][slot]; cur[slot] is live.
  // Slots A_PAI and A_TETAV are always double.
  mlc::Buf adv[2][mlc::kMaxSlots];
  int cur[mlc::kMaxSlots] = {};
  mlc::Buf tke, tketen, chiten;  // ibltyp 2: kz+1 levels; ntr > 0: ntr blocks
  // statics (bulk type)
  mlc::Buf fmz, fmzf, rfmzu, rfmzv, bdywtu, bdywtv, bdywtw, zeta, dz;
  mlc::Buf mx, mu, mv, hx, hy, coru, corv, mx2, rmx, rmu, rmv, tanrot;
  double *fmz64 = nullptr, *fmzf64 = nullptr;   // double metrics of the solve
  // surface geopotential ht | htu | htv ([3][iy][pj]) and the per-level
  // gzitap | bzitap tables (2 (2 kz + 1)) of the formed terrain metrics
  double *hgeo = nullptr, *ztab = nullptr;
  double *geo[3] = {};           // the three planes of hgeo
  double mo_h = 0.0;             // MLC_PARAM_MO_H
  bool zmet_ok = false;          // mlc_init found the formed metrics identical
  // boundary weights u, v, w as column ids [3][iy][pj] and column tables
  int *bidx = nullptr;
  mlc::Buf btab[3];
  int bcols[3] = {};
  double *vec1d = nullptr;       // gzitak | gzitakh | xkdamp | xknu | ffilt
  // scratch
  mlc::Buf s, zdiv2, zrfmu, zrfmd;
  mlc::Buf wz2[mlc::kMaxSlots];
  double *tetavf = nullptr;      // always double
  double *colw = nullptr;        // streamed column solve: wt | wkk | z2 (double)
  double *red = nullptr;         // reduction partials
  int device = -1;               // the device current at create
  size_t red_len = 0;

  // pinned staging for whole-plane transfers
  void *pinned = nullptr;
  size_t pinned_len = 0;

  // host copies of the small static inputs
  std::vector<double> h_mx, h_mu, h_mv, h_rlat, h_ffilt;
  bool have_static[64] = {};

  // timing
  bool prof = false;
  unsigned prof_mask = 0;          // kernel slots bracketed by events
  struct Ev { int slot; hipEvent_t a, b; };
  std::vector<Ev> pending;
  std::vector<hipEvent_t> pool;
  long long launches[mlc::K_COUNT] = {};
  double ms[mlc::K_COUNT] = {};

  // lateral-boundary data of a real-case run (double in both forms)
  struct Lbc {
    bool on = false;
    mlc_lbc_desc desc{};
    double *ps0 = nullptr, *ps1 = nullptr, *u0 = nullptr, *u1 = nullptr, *v0 = nullptr,
           *v1 = nullptr, *t0 = nullptr, *t1 = nullptr, *pai0 = nullptr, *pai1 = nullptr,
           *qv0 = nullptr, *qv1 = nullptr, *qc0 = nullptr, *qc1 = nullptr, *qi0 = nullptr,
           *qi1 = nullptr, *xfu = nullptr, *xfv = nullptr, *xfc = nullptr, *tnudge = nullptr;
    // spectral nudging: bases, profile, scratch, accumulated time
    double *bvx = nullptr, *bvy = nullptr, *cnudge = nullptr, *zn = nullptr, *sx = nullptr,
           *sy = nullptr;
    bool have_spectral = false;
    double tspectral = 0.0;
  } lbc;

  // physics input slices (mkslice), double
  struct Slice {
    double *zetaf = nullptr, *xlat = nullptr;
    double *pf = nullptr, *th = nullptr, *rh = nullptr, *omega = nullptr, *rhox = nullptr,
           *tp = nullptr, *th700 = nullptr, *ptrop = nullptr, *ktrop = nullptr,
           *kmxpbl = nullptr;
    bool on = false, have_zetaf = false, have_xlat = false;
  } sl;

  // launch-shape choices (autotuned per domain shape)
  struct Tune {
    int waf_h = 1;               // horizontal WAF: 0 row-streaming, 1 LDS tile
    int waf_rows = 16;           // row-streaming WAF: i rows per wavefront
    int waf_group = 4;           // row-streaming WAF: fields per lane (2 or 4)
    int level_chunk = 8;         // divergence/momentum: levels per thread
    int mom_chunk = 0;           // momentum, if not 0: levels per thread
    int stag_chunk = 1;          // staggering passes: levels per thread
    int column = 2;              // column solve: 0 LDS three-phase, 1 streamed, 2 chunked
    int wafv_group = 4;          // vertical WAF: fields per thread (1, 2 or 4)
    int xcd_map = 1;             // XCD-contiguous tile order of single-domain launches
    int waf_tile = 4;            // LDS-tile WAF: output rows per tile (4 or 8; FP64 starts at 8)
    int waf_hgroup = 0;          // LDS-tile WAF: fields per block (0 = the whole batch)
    int column_j = 32;           // chunked column solve: columns per block (16 or 32)
    int mixed_merge = 1;         // mixed form: float and double WAF batches in one launch
                                 // (bit 0 vertical, bit 1 horizontal)
    int terrain = 0;             // stages forming the terrain metrics from the orography
                                 // (mask: 1 divergence, 2 column, 4 staggering, 8 WAF-H)
  } tune;

  // while non-null, stages are appended here instead of launched
  std::vector<mlc::LaunchRec> *rec = nullptr;
  bool batch_member = false;     // owned by an mlc_batch
  bool own_stream = true;

  const mlc::Buf &cur_adv(int slot) const { return adv[cur[slot]][slot]; }
  const mlc::Buf &alt_adv(int slot) const { return adv[1 - cur[slot]][slot]; }
};

namespace mlc {

// Error reporting (thread-local message).
int fail(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);
#define MLC_HIP(call)                                                  \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) return ::mlc::hip_check(e_, #call);          \
  } while (0)

// Handle lifecycle; `shared` (may be null) is a stream the handle uses
// instead of creating its own.
int create_handle(const mlc_grid_desc *g, hipStream_t shared, mlc_handle **out);
void destroy_handle(mlc_handle *h);
int autotune(mlc_handle *h, int nsteps, double *best_ms);

// A device copy of the state a step reads (and the ping-pong indices), for
// timing trials that must leave the state as it was.  take / restore are
// asynchronous on the handle's stream.
struct Snapshot {
  char *dev = nullptr;
  int cur[kMaxSlots] = {};
  bool dirty = false;
};
int snapshot_take(mlc_handle *h, Snapshot *snap);
int snapshot_restore(mlc_handle *h, const Snapshot &snap);
void snapshot_free(Snapshot *snap);

// Timing brackets around one launch on h->stream.
// Per-launch timing: a profiled launch of `slot` gets two events that the
// launch itself records at the kernel's start and end (hipExtLaunchKernel),
// so the time is the kernel's alone, without the dispatch gap that events
// recorded around it would add.
void prof_events(mlc_handle *h, int slot, hipEvent_t *a, hipEvent_t *b);
void prof_push(mlc_handle *h, int slot, hipEvent_t a, hipEvent_t b);

// Launchers (mlc_kernels.hip).  Each returns 0 or an MLC_E_* code and
// dispatches on the handle's form.
int launch_sound(mlc_handle *h);   // one sound() call: mo_nsound sub-steps
int launch_advection(mlc_handle *h);
int launch_dyn_tail(mlc_handle *h);
int launch_diagnostics(mlc_handle *h);
int launch_massck_noalloc(mlc_handle *h, double out[4]);
int launch_init_restart(mlc_handle *h);
// the dycore tail and the diagnostics in one pass (they are pointwise and
// adjacent when no boundary stage runs between them)
int launch_dyn_tail_diagnostics(mlc_handle *h);
// One physics-off step after the tendency reset: nadv x (sound + advection),
// dycore tail + diagnostics, surface pressure, status_update.
int step_body(mlc_handle *h);
int launch_status_update(mlc_handle *h, double dtinc);
// dycore tail + diagnostics + surface pressure + status_update of a step
// whose tendencies are zero (physics off)
int launch_step_tail(mlc_handle *h);
int launch_boundary(mlc_handle *h, double x1);
int launch_mkslice(mlc_handle *h, const mlc_slice_desc &s);
int launch_fill(mlc_handle *h, const Buf &b, long long n, double value);
int launch_fill64(mlc_handle *h, double *p, long long n, double value);
int launch_waf_metrics(mlc_handle *h);
int launch_terrain_check(mlc_handle *h, long long *nbad);
int launch_massck(mlc_handle *h, double out[4]);
int launch_check_finite(mlc_handle *h, double *psmax, double *psmin);

}  // namespace mlc
