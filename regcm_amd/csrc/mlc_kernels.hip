// Kernels of the MOLOCH step for gfx950, in two forms.
//
// Every kernel is a template on the bulk real type R.  With R = double (the
// MLC_FP64 form) each kernel reproduces the reference's per-point
// arithmetic operation for operation (Fortran left-to-right association,
// explicit parentheses kept); the file is compiled with -ffp-contract=off
// so no multiply-add is fused and that form is bit-reproducible against a
// CPU build of the same recurrences.  With R = float (MLC_MIXED) the
// stencil bulk runs in float; the Exner function, theta_v, w, tetavf, the
// column solve, the Exner update, the momentum arithmetic, the advection of
// pai and theta_v, p and the budget sums stay in double.  Fusion only moves work between loop nests whose inputs are not
// modified in between; it never changes an operand.
//
// Launch shapes: 64x4 thread blocks over (j, i) tiles, j on the wavefront
// lanes so that every global access of a wavefront is one contiguous row
// segment; grid.z runs over levels or level chunks.  The column solve runs
// its recurrences one column per lane from LDS.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include <hip/hip_ext.h>

#include "mlc_internal.h"

namespace mlc {

const char *const kKernelNames[K_COUNT] = {
    "sound_div", "sound_column", "sound_momentum", "adv_stagger_in",
    "waf_vertical", "waf_horizontal", "adv_curvature", "adv_stagger_out",
    "diag_ps", "tail", "uv_stagger", "fill", "frame", "massck", "finite", "lbc",
    "mkslice"};

constexpr int TJ = 64;   // j extent of a tile (wavefront width)
constexpr int TI = 4;    // i extent of a tile

template <typename T>
struct AdvPtrs {
  const T *src[kMaxAdv];
  T *dst[kMaxAdv];
};

template <typename R>
struct QPtrs {
  R *q[10];
  const R *qt[10];
};

template <typename R>
__device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }

// local_flow_param, Main/mod_moloch.F90:1571-1590: minden = 1e-30 and
// minnum = real(4) rounding of 1e-30 in double; the single-precision build
// of the reference uses 1e-15 for both (:1574-1575).
template <typename R> struct FlowLim;
template <> struct FlowLim<double> {
  static __device__ __forceinline__ double den() { return 1.0e-30; }
  static __device__ __forceinline__ double num() { return (double)1.0e-30f; }
};
template <> struct FlowLim<float> {
  static __device__ __forceinline__ float den() { return 1.0e-15f; }
  static __device__ __forceinline__ float num() { return 1.0e-15f; }
};

template <typename R>
__device__ __forceinline__ R flow_param(R num, R den) {
  if (fabs(den) < FlowLim<R>::den())
    return fabs(num) < FlowLim<R>::num() ? R(1) : R(0);
  return num / den;
}

// b = max(0, min(2, max(r, min(2r, 1))))  (Main/mod_moloch.F90:882)
template <typename R>
__device__ __forceinline__ R waf_b(R r) {
  return rmax(R(0), rmin(R(2), rmax(r, rmin(R(2) * r, R(1)))));
}

// ---------------------------------------------------------------------------
// Entry points.  Every stage is an argument struct A with a device body
// A::run(a, bx, by, bz).  k_one launches it on one domain with the block
// index as tile index; k_many launches it on a batch of domains in one grid:
// block b runs tile (map[b].y, map[b].z, map[b].w) of domain map[b].x with
// that domain's arguments tab[map[b].x].
//
// MinWaves<A> is the occupancy a stage asks of the register allocator
// (waves per SIMD): a grid whose waves just miss one round of the chip at
// the default allocation runs a second, mostly idle round.
template <class A> struct MinWaves { static constexpr int value = 1; };

template <class A>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<A>::value)))
void k_one(A a) {
  A::run(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Stages whose tiles share horizontal halos take the XCD-aware order below;
// column-streaming and pointwise stages keep the plain order.
template <class A> struct XcdOrder { static constexpr bool value = true; };

// The same with an XCD-aware tile order: the hardware hands workgroup b of a
// launch to XCD b mod 8, so XCD x is given the contiguous run of tiles
// [start(x), start(x) + count(x)) in x-fastest order, and neighbouring
// tiles, whose halos overlap, meet in the same L2.
template <class A>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<A>::value)))
void k_one_xcd(A a) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned x = b % 8u, p = b / 8u;
  const unsigned q = nb / 8u, r = nb % 8u;
  const unsigned t = x * q + (x < r ? x : r) + p;
  A::run(a, (int)(t % gx), (int)((t / gx) % gy), (int)(t / (gx * gy)));
}

template <class A>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves<A>::value)))
void k_many(const A *__restrict__ tab, const uint4 *__restrict__ map) {
  const uint4 m = map[blockIdx.x];
  A::run(tab[m.x], (int)m.y, (int)m.z, (int)m.w);
}

// Marks loaded values as used here: the compiler waits for their loads at
// this point (together), and can neither sink a load into a later branch
// nor issue it behind a store that precedes its use.  An empty statement;
// it emits no instruction.
template <class X>
__device__ __forceinline__ void issued_one(const X &x) {
  asm volatile("" ::"v"(x));
}
template <class... X>
__device__ __forceinline__ void issued(const X &...x) {
  (issued_one(x), ...);
}

// ---------------------------------------------------------------------------
// Acoustic sub-step, part 1: horizontal divergence zdiv2 on every cross point
// (Main/mod_moloch.F90:602-618) and the generalized vertical velocity s on
// interior columns (:582-597), with w(kz+1) = -s(kz+1).  Each thread owns a
// chunk of levels [k0, k0+kc) of one column: the 2-D metrics are loaded once
// and u, v of level k-1 are carried in registers.
template <typename R>
struct SoundDiv {
  Dims d;
  const R * u;
  const R * v;
  const R * rfmzu;
  const R * rfmzv;
  const R * fmz;
  const R * mx;
  const R * mx2;
  const R * rmu;
  const R * rmv;
  const R * hx;
  const R * hy;
  const double * gzitak;
  R * zdiv2;
  R * s;
  double * w;
  double dtrdx_;
  double dtrdy_;
  int rotllr;
  int kc;
  ZMet zm;
  static __device__ __forceinline__ void run(const SoundDiv &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SoundDiv<R>::run(const SoundDiv &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const R *__restrict__ u = args_.u;
  const R *__restrict__ v = args_.v;
  const R *__restrict__ rfmzu = args_.rfmzu;
  const R *__restrict__ rfmzv = args_.rfmzv;
  const R *__restrict__ fmz = args_.fmz;
  const R *__restrict__ mx = args_.mx;
  const R *__restrict__ mx2 = args_.mx2;
  const R *__restrict__ rmu = args_.rmu;
  const R *__restrict__ rmv = args_.rmv;
  const R *__restrict__ hx = args_.hx;
  const R *__restrict__ hy = args_.hy;
  const double *__restrict__ gzitak = args_.gzitak;
  R *__restrict__ zdiv2 = args_.zdiv2;
  R *__restrict__ s = args_.s;
  double *__restrict__ w = args_.w;
  const double dtrdx_ = args_.dtrdx_;
  const double dtrdy_ = args_.dtrdy_;
  const int rotllr = args_.rotllr;
  const int kc = args_.kc;
  const ZMet zm = args_.zm;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int k0 = 1 + (int)bz * kc;
  const int k1 = k0 + kc - 1 < d.kz ? k0 + kc - 1 : d.kz;
  if (j > d.je2 || i > d.ie2) return;
  const R dtrdx = R(dtrdx_), dtrdy = R(dtrdy_);
  const long long P = d.plane;
  const long long e = d.oj(j, 1), n = d.oi(i, 1);   // east / north neighbours
  const long long c2 = d.at2(j, i);
  const R rmua = rmu[c2], rmub = rmu[c2 + e];
  const R rmva = rmv[c2], rmvb = rmv[c2 + n];
  const R mfac = rotllr ? mx[c2] : mx2[c2];
  const bool interior = j >= d.jc1 && j <= d.jc2 && i >= d.ic1 && i <= d.ic2;
  R hxa = 0, hxb = 0, hya = 0, hyb = 0;
  R ua = 0, ub = 0, va = 0, vb = 0;      // u, v of level k-1
  if (interior) {
    hxa = hx[c2]; hxb = hx[c2 + e];
    hya = hy[c2]; hyb = hy[c2 + n];
    if (k0 >= 2) {
      const long long cm = d.at(j, i, k0 - 1);
      ua = u[cm]; ub = u[cm + e]; va = v[cm]; vb = v[cm + n];
    }
  }
  // the level loop, with the metrics formed from the orography (ZC) or
  // streamed: one loop per case, so that neither pays for the other
  auto levels = [&](auto zc) __attribute__((always_inline)) {
  for (int k = k0; k <= k1; ++k) {
    const long long c = d.at(j, i, k);
    const R u0 = u[c], u1 = u[c + e], v0 = v[c], v1 = v[c + n];
    // loaded with the level's other operands: a load issued after the zdiv2
    // store would wait for that store too (one counter covers both)
    const R gzk = R(gzitak[k - 1]);
    R ru0, ru1, rv0, rv1, fz;
    if (decltype(zc)::value) {
      ru0 = R(zm.rfmzu(c2, k)); ru1 = R(zm.rfmzu(c2 + e, k));
      rv0 = R(zm.rfmzv(c2, k)); rv1 = R(zm.rfmzv(c2 + n, k));
      fz = R(zm.fmz(c2, k));
    } else {
      ru0 = rfmzu[c]; ru1 = rfmzu[c + e]; rv0 = rfmzv[c]; rv1 = rfmzv[c + n];
      fz = fmz[c];
    }
    R zum, zup;
    if (rotllr) {
      zum = dtrdx * u0 * ru0;
      zup = dtrdx * u1 * ru1;
    } else {
      zum = dtrdx * u0 * ru0 * rmua;
      zup = dtrdx * u1 * ru1 * rmub;
    }
    const R zvm = dtrdy * v0 * rv0 * rmva;
    const R zvp = dtrdy * v1 * rv1 * rmvb;
    zdiv2[c] = fz * mfac * ((zup - zum) + (zvp - zvm));
    if (interior) {
      if (k >= 2) {
        const R zuh = (u0 + ua) * hxa + (u1 + ub) * hxb;
        const R zvh = (v0 + va) * hya + (v1 + vb) * hyb;
        s[c] = R(-0.25) * (zuh + zvh) * gzk;
      }
      if (k == d.kz) {
        const R zuh = u0 * hxa + u1 * hxb;
        const R zvh = v0 * hya + v1 * hyb;
        const R sk = R(-0.5) * (zuh + zvh);
        s[c + P] = sk;
        w[c + P] = -(double)sk;
      }
    }
    ua = u0; ub = u1; va = v0; vb = v1;
  }
  };
  if (zm.on)
    levels(std::true_type{});
  else
    levels(std::false_type{});
}

// ---------------------------------------------------------------------------
// Acoustic sub-step, part 2: divergence filter (:531-543), vertical-
// divergence term (:627-630), the implicit w solve (:634-664) and the Exner
// update (:668-671) for a row of kColJ interior columns, in double.
// Phase 1 (all threads, levels spread over threadIdx.y): the filtered
// divergence zd2(k), tetavf, the explicit w and the tridiagonal coefficients
// go to LDS.  Phase 2 (one lane per column): only the two recurrences of the
// elimination, from LDS.  Phase 3 (all threads): w and the Exner update.
// zdiv2 is read-only here, so the 5-point Laplacian always sees unfiltered
// neighbours, as the reference's separate `laplacian` buffer guarantees.  On
// the first sub-step of a sound() call tetavf is formed from tetav
// (:564-566) instead of loaded.
constexpr int kColJ = 32;   // columns per block (along j)
constexpr int kColK = 8;    // threads per column (levels spread over them)

template <typename R>
struct SoundColumn {
  Dims d;
  const R * zdiv2;
  const R * s;
  const double * tetav;
  const double * fmz;
  const double * fmzf;
  BdyWt<R> bdw;
  const double * xknu;
  const double * ffilt;
  double * tetavf;
  double * w;
  double * pai;
  double dts;
  double dtrdz;
  double zcs2;
  int divfilter;
  int first;
  static __device__ __forceinline__ void run(const SoundColumn &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SoundColumn<R>::run(const SoundColumn &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const R *__restrict__ zdiv2 = args_.zdiv2;
  const R *__restrict__ s = args_.s;
  const double *__restrict__ tetav = args_.tetav;
  const double *__restrict__ fmz = args_.fmz;
  const double *__restrict__ fmzf = args_.fmzf;
  const BdyWt<R> bdw = args_.bdw;
  const double *__restrict__ xknu = args_.xknu;
  const double *__restrict__ ffilt = args_.ffilt;
  double *__restrict__ tetavf = args_.tetavf;
  double *__restrict__ w = args_.w;
  double *__restrict__ pai = args_.pai;
  const double dts = args_.dts;
  const double dtrdz = args_.dtrdz;
  const double zcs2 = args_.zcs2;
  const int divfilter = args_.divfilter;
  const int first = args_.first;

  extern __shared__ double smem[];
  const int kz = d.kz;
  const int L = (kz + 1) * kColJ;
  double *s_zu = smem, *s_zd = smem + L, *s_rh = smem + 2 * L,
         *s_z2 = smem + 3 * L;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int j = d.jc1 + bx * kColJ + tx;
  const int i = d.ic1 + by;
  const bool ok = j <= d.jc2;
  const long long P = d.plane;
  const int jj = ok ? j : d.jc1;
  const long long c1 = d.at(jj, i, 1);
  const long long c2 = d.at2(jj, i);
  const long long ow = d.oj(jj, -1), oe = d.oj(jj, 1), os = d.oi(i, -1), on = d.oi(i, 1);

  if (ok) {
    for (int k = 1 + ty; k <= kz; k += kColK) {
      const long long c = c1 + (long long)(k - 1) * P;
      double z = zdiv2[c];
      if (divfilter) {
        const double lap = ((double)zdiv2[c + ow] + (double)zdiv2[c + oe] +
                            (double)zdiv2[c + os] + (double)zdiv2[c + on] -
                            4.0 * z);
        z = z + xknu[k - 1] * lap;
      }
      s_z2[k * kColJ + tx] = z + (double)bdw.at(c2, k) * dtrdz * fmz[c] *
                                     ((double)s[c] - (double)s[c + P]);
    }
  }
  __syncthreads();
  if (ok) {
    for (int k = 2 + ty; k <= kz; k += kColK) {
      const long long c = c1 + (long long)(k - 1) * P, cm = c - P;
      const double tvk = tetav[c], tvkm1 = tetav[cm];
      double tvf = first ? 0.5 * (tvkm1 + tvk) : tetavf[c];
      const double wk = w[c], fzf = fmzf[c];
      tvf = tvf - wk * fzf * dtrdz * (tvkm1 - tvk);
      tetavf[c] = tvf;
      const double zrom1w = kCpd * tvf * fzf;
      const double pk = pai[c], pkm1 = pai[cm];
      const double zkm1 = s_z2[(k - 1) * kColJ + tx], zk = s_z2[k * kColJ + tx];
      double zwexpl = wk - zrom1w * dtrdz * (pkm1 - pk) - kEgrav * dts;
      zwexpl = zwexpl + kRdrcv * zrom1w * dtrdz * (pkm1 * zkm1 - pk * zk);
      s_zu[k * kColJ + tx] = zcs2 * fmz[cm] * zrom1w * pkm1 + ffilt[k - 1];
      s_zd[k * kColJ + tx] = zcs2 * fmz[c] * zrom1w * pk + ffilt[k - 1];
      s_rh[k * kColJ + tx] = zwexpl;
    }
  }
  __syncthreads();
  const double w1 = ok ? w[c1] : 0.0;                              // w(1)
  const double wtop = ok ? w[c1 + (long long)kz * P] : 0.0;        // w(kz+1)
  if (ok && ty == 0) {
    double wk_up = 0.0, w_up = wtop;                               // wwkw(kz+1) = 0
    for (int k = kz; k >= 2; --k) {
      const double zu = s_zu[k * kColJ + tx], zd = s_zd[k * kColJ + tx];
      const double zrapp = 1.0 / (1.0 + zd + zu - zd * wk_up);
      const double wt = zrapp * (s_rh[k * kColJ + tx] + zd * w_up);
      const double wkk = zrapp * zu;
      s_rh[k * kColJ + tx] = wt;
      s_zu[k * kColJ + tx] = wkk;
      w_up = wt;
      wk_up = wkk;
    }
    double below = w1;
    for (int k = 2; k <= kz; ++k) {
      const double wf = s_rh[k * kColJ + tx] + s_zu[k * kColJ + tx] * below;
      s_rh[k * kColJ + tx] = wf;
      below = wf;
    }
  }
  __syncthreads();
  if (ok) {
    for (int k = 1 + ty; k <= kz; k += kColK) {
      const long long c = c1 + (long long)(k - 1) * P;
      const double wk = (k == 1) ? w1 : s_rh[k * kColJ + tx];
      const double wkp1 = (k == kz) ? wtop : s_rh[(k + 1) * kColJ + tx];
      if (k >= 2) w[c] = wk;
      pai[c] = pai[c] * (1.0 - kRdrcv * (s_z2[k * kColJ + tx] +
                                         (dtrdz * fmz[c] * (wk - wkp1))));
    }
  }
}

// The same column stage with each thread owning a contiguous chunk of at
// most CC levels of its column: the filtered divergence z2 of its levels
// (and of the level above its chunk) stays in registers, so LDS holds only
// the three tridiagonal arrays (zu, zd, rhs) and the stage needs two
// barriers.  CJ columns per block (along j), CK threads per column, at
// most CC levels per thread; requires kz <= CK * CC, and runs with the
// smallest CC that holds the column (6 or 3 for kz <= 48, else 8) so that
// no thread carries unused levels.  Same operations as SoundColumn.
constexpr int kColC = 8;

// LDS of a chunked column block of cj columns: the three tridiagonal arrays
// of levels 2..kz and the interleaved gzitap / bzitap tables (2 kz + 1
// levels each) of the formed terrain metrics
inline size_t col_chunk_lds(int kz, int cj) {
  return (3 * (size_t)(kz - 1) * cj + 2 * (size_t)(2 * kz + 1)) * sizeof(double);
}

template <typename R, int CJ, int CK, int CC = kColC>
struct SoundColumnChunk {
  Dims d;
  const R *zdiv2, *s;
  BdyWt<R> bdw;
  const double *tetav, *fmz, *fmzf, *xknu, *ffilt;
  double *tetavf, *w, *pai;
  double dts, dtrdz, zcs2;
  int divfilter, first;
  ZMet zm;
  static __device__ __forceinline__ void run(const SoundColumnChunk &args_, int bx, int by, int bz);
};

template <typename R, int CJ, int CK, int CC>
__device__ __forceinline__ void SoundColumnChunk<R, CJ, CK, CC>::run(
    const SoundColumnChunk &args_, int bx, int by, int bz) {
  extern __shared__ double smem[];
  const Dims d = args_.d;
  const R *__restrict__ zdiv2 = args_.zdiv2;
  const R *__restrict__ s = args_.s;
  const BdyWt<R> bdw = args_.bdw;
  const double *__restrict__ tetav = args_.tetav;
  const double *__restrict__ fmz = args_.fmz;
  const double *__restrict__ fmzf = args_.fmzf;
  const double *__restrict__ xknu = args_.xknu;
  const double *__restrict__ ffilt = args_.ffilt;
  double *__restrict__ tetavf = args_.tetavf;
  double *__restrict__ w = args_.w;
  double *__restrict__ pai = args_.pai;
  const double dts = args_.dts, dtrdz = args_.dtrdz, zcs2 = args_.zcs2;
  const int divfilter = args_.divfilter, first = args_.first;
  const ZMet zm = args_.zm;
  const int kz = d.kz;
  // levels 2..kz of the tridiagonal arrays, level k at row k - 2; then,
  // with the formed metrics, gzitap and bzitap of every level, interleaved
  const int L = (kz - 1) * CJ;
  double *s_zu = smem, *s_zd = smem + L, *s_rh = smem + 2 * L;
  double *s_gb = smem + 3 * L;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int j = d.jc1 + bx * CJ + tx;
  const int i = d.ic1 + by;
  const bool ok = j <= d.jc2;
  const long long P = d.plane;
  const int jj = ok ? j : d.jc1;
  const long long c1 = d.at(jj, i, 1);
  const long long c2 = d.at2(jj, i);
  const long long ow = d.oj(jj, -1), oe = d.oj(jj, 1), os = d.oi(i, -1), on = d.oi(i, 1);
  const int kc = (kz + CK - 1) / CK;
  const int ka = 1 + ty * kc;                       // first level of the chunk
  const int kb = ka + kc - 1 < kz ? ka + kc - 1 : kz;
  // w(1) and w(kz+1) are not written by this stage; their loads are issued
  // first so that their latency overlaps phase 1.  Lanes past the domain
  // (!ok) load from the valid column c1 and store nothing.
  const double w1 = w[c1];                                         // w(1)
  const double wtop = w[c1 + (long long)kz * P];                   // w(kz+1)
  // The formed metrics' per-level tables are read from LDS and the
  // orography term ht(j,i)*regrav is loaded once: read from global memory
  // at every level, each division waited for its own table loads (and ht
  // was loaded again after every tetavf store, which could alias it).
  // the boundary-weight column (its index load, on which every weight's
  // address depends, goes out before the tables' barrier)
  const R *__restrict__ bcol = bdw.col(c2);
  double hreg = 0.0;
  if (zm.on) {
    hreg = zm.hc[c2] * kRegrav;
    for (int t = ty * CJ + tx; t < 2 * kz + 1; t += CJ * CK) {
      s_gb[2 * t] = zm.ga[t];
      s_gb[2 * t + 1] = zm.ba[t];
    }
    __syncthreads();
  }
  // ZMet::fmz(c2, k) = 1 / den(t = k - 1), fmzf(c2, k) = 1 / den(t = kz + k - 1)
  auto zden = [&](int t) { return s_gb[2 * t] * hreg + s_gb[2 * t + 1]; };
  // z2r[m], pr[m], fr[m] = z2, pai, fmz of level ka - 1 + m, m = 0 .. kc;
  // pai and fmz of the chunk stay in registers for the Exner update.
  // Phase 1 is branch-free up to its stores: every load reads a level
  // clamped into the column and the unused values are masked, so the loads
  // of all levels can be in flight together.
  double z2r[CC + 1], pr[CC + 1], fr[CC + 1];
  // phase 1 with the metrics formed from the orography (ZC) or streamed
  auto phase1 = [&](auto zc) __attribute__((always_inline)) {
    constexpr bool ZC = decltype(zc)::value;
    auto z2_levels = [&](auto df) {
#pragma unroll
      for (int m = 0; m <= CC; ++m) {
        const int k = ka - 1 + m;
        const bool in = m <= kb - ka + 1 && k >= 1;
        const int kk = k < 1 ? 1 : (k > kz ? kz : k);
        const long long c = c1 + (long long)(kk - 1) * P;
        const double fzk = ZC ? 1.0 / zden(kk - 1) : fmz[c], pk = pai[c];
        double z = zdiv2[c];
        if (decltype(df)::value) {
          const double lap = ((double)zdiv2[c + ow] + (double)zdiv2[c + oe] +
                              (double)zdiv2[c + os] + (double)zdiv2[c + on] - 4.0 * z);
          z = z + xknu[kk - 1] * lap;
        }
        z = z + (double)bcol[kk - 1] * dtrdz * fzk * ((double)s[c] - (double)s[c + P]);
        fr[m] = in ? fzk : 0.0;
        pr[m] = in ? pk : 0.0;
        z2r[m] = in ? z : 0.0;
      }
    };
    if (divfilter)
      z2_levels(std::true_type{});
    else
      z2_levels(std::false_type{});
    // the tridiagonal coefficients of levels ka .. kb (>= 2); the inputs of
    // the next level are loaded before this level's stores
    auto lev = [&](int m) {
      const int k = ka - 1 + m;
      return k < 2 ? 2 : (k > kz ? kz : k);
    };
    double q_tv, q_tm, q_tf, q_w, q_fz, q_ff;
    auto fetch = [&](int m) {
      const long long c = c1 + (long long)(lev(m) - 1) * P;
      q_tv = tetav[c];
      q_tm = tetav[c - P];
      q_tf = tetavf[c];
      q_w = w[c];
      q_fz = ZC ? 1.0 / zden(kz + lev(m) - 1) : fmzf[c];
      q_ff = ffilt[lev(m) - 1];
    };
    fetch(1);
#pragma unroll
    for (int m = 1; m <= CC; ++m) {
      const double tvk = q_tv, tvkm1 = q_tm, tvf0 = q_tf, wk = q_w, fzf = q_fz, ffk = q_ff;
      if (m < CC) fetch(m + 1);
      const int k = ka - 1 + m;
      double tvf = first ? 0.5 * (tvkm1 + tvk) : tvf0;
      tvf = tvf - wk * fzf * dtrdz * (tvkm1 - tvk);
      const double zrom1w = kCpd * tvf * fzf;
      const double pk = pr[m], pkm1 = pr[m - 1];
      const double zkm1 = z2r[m - 1], zk = z2r[m];
      double zwexpl = wk - zrom1w * dtrdz * (pkm1 - pk) - kEgrav * dts;
      zwexpl = zwexpl + kRdrcv * zrom1w * dtrdz * (pkm1 * zkm1 - pk * zk);
      const double zu = zcs2 * fr[m - 1] * zrom1w * pkm1 + ffk;
      const double zd = zcs2 * fr[m] * zrom1w * pk + ffk;
      if (ok && k <= kb && k >= 2) {
        tetavf[c1 + (long long)(k - 1) * P] = tvf;
        s_zu[(k - 2) * CJ + tx] = zu;
        s_zd[(k - 2) * CJ + tx] = zd;
        s_rh[(k - 2) * CJ + tx] = zwexpl;
      }
    }
  };
  if (ka <= kz) {
    if (zm.on)
      phase1(std::true_type{});
    else
      phase1(std::false_type{});
  }
  __syncthreads();
  if (ok && ty == 0) {
    double wk_up = 0.0, w_up = wtop;                               // wwkw(kz+1) = 0
    for (int k = kz; k >= 2; --k) {
      const double zu = s_zu[(k - 2) * CJ + tx], zd = s_zd[(k - 2) * CJ + tx];
      const double zrapp = 1.0 / (1.0 + zd + zu - zd * wk_up);
      const double wt = zrapp * (s_rh[(k - 2) * CJ + tx] + zd * w_up);
      const double wkk = zrapp * zu;
      s_rh[(k - 2) * CJ + tx] = wt;
      s_zu[(k - 2) * CJ + tx] = wkk;
      w_up = wt;
      wk_up = wkk;
    }
    double below = w1;
    for (int k = 2; k <= kz; ++k) {
      const double wf = s_rh[(k - 2) * CJ + tx] + s_zu[(k - 2) * CJ + tx] * below;
      s_rh[(k - 2) * CJ + tx] = wf;
      below = wf;
    }
  }
  __syncthreads();
  if (ok && ka <= kz) {
#pragma unroll
    for (int m = 1; m <= CC; ++m) {
      const int k = ka - 1 + m;
      if (k <= kb) {
        const long long c = c1 + (long long)(k - 1) * P;
        const double wk = (k == 1) ? w1 : s_rh[(k - 2) * CJ + tx];
        const double wkp1 = (k == kz) ? wtop : s_rh[(k - 1) * CJ + tx];
        if (k >= 2) w[c] = wk;
        pai[c] = pr[m] * (1.0 - kRdrcv * (z2r[m] + (dtrdz * fr[m] * (wk - wkp1))));
      }
    }
  }
}

// The same column stage, streamed one column per lane with no LDS: the
// forward elimination runs k = kz .. 2 carrying level k's values in
// registers and writes the elimination results wt, wkk and the filtered
// divergence z2 to scratch arrays; the back substitution k = 2 .. kz reads
// them back (most recent first) and applies the Exner update.  Every value
// is formed with the operations of SoundColumn.
template <typename R>
struct SoundColumnStream {
  Dims d;
  const R *zdiv2, *s;
  BdyWt<R> bdw;
  const double *tetav, *fmz, *fmzf, *xknu, *ffilt;
  double *tetavf, *w, *pai, *wtb, *wkb, *z2b;
  double dts, dtrdz, zcs2;
  int divfilter, first;
  static __device__ __forceinline__ void run(const SoundColumnStream &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SoundColumnStream<R>::run(const SoundColumnStream &args_, int bx,
                                                          int by, int bz) {
  const Dims d = args_.d;
  const int j = d.jc1 + bx * TJ + threadIdx.x;
  const int i = d.ic1 + by * TI + threadIdx.y;
  if (j > d.jc2 || i > d.ic2) return;
  const R *__restrict__ zdiv2 = args_.zdiv2;
  const R *__restrict__ s = args_.s;
  const BdyWt<R> bdw = args_.bdw;
  const double *__restrict__ tetav = args_.tetav;
  const double *__restrict__ fmz = args_.fmz;
  const double *__restrict__ fmzf = args_.fmzf;
  const double *__restrict__ xknu = args_.xknu;
  const double *__restrict__ ffilt = args_.ffilt;
  double *__restrict__ tetavf = args_.tetavf;
  double *__restrict__ w = args_.w;
  double *__restrict__ pai = args_.pai;
  double *__restrict__ wtb = args_.wtb;
  double *__restrict__ wkb = args_.wkb;
  double *__restrict__ z2b = args_.z2b;
  const double dts = args_.dts, dtrdz = args_.dtrdz, zcs2 = args_.zcs2;
  const int divfilter = args_.divfilter, first = args_.first;
  const int kz = d.kz;
  const long long P = d.plane;
  const long long c1 = d.at(j, i, 1);
  const long long c2 = d.at2(j, i);
  const long long ow = d.oj(j, -1), oe = d.oj(j, 1), os = d.oi(i, -1), on = d.oi(i, 1);
  // filtered divergence plus the vertical-divergence term at level k
  auto z2_of = [&](long long c, int k, double sk, double skp1) -> double {
    double z = zdiv2[c];
    if (divfilter) {
      const double lap = ((double)zdiv2[c + ow] + (double)zdiv2[c + oe] +
                          (double)zdiv2[c + os] + (double)zdiv2[c + on] - 4.0 * z);
      z = z + xknu[k - 1] * lap;
    }
    return z + (double)bdw.at(c2, k) * dtrdz * fmz[c] * (sk - skp1);
  };
  const long long ckz = c1 + (long long)(kz - 1) * P;
  const double w1 = w[c1];                  // w(1)
  const double wtop = w[ckz + P];           // w(kz+1)
  double s_k = s[ckz];
  double z_k = z2_of(ckz, kz, s_k, (double)s[ckz + P]);
  z2b[ckz] = z_k;
  double tv_k = tetav[ckz], pai_k = pai[ckz], fmz_k = fmz[ckz];
  double wk_up = 0.0, w_up = wtop;          // wwkw(kz+1) = 0
  for (int k = kz; k >= 2; --k) {
    const long long c = c1 + (long long)(k - 1) * P, cm = c - P;
    const double s_km1 = s[cm];
    const double z_km1 = z2_of(cm, k - 1, s_km1, s_k);
    z2b[cm] = z_km1;
    const double tv_km1 = tetav[cm], pai_km1 = pai[cm], fmz_km1 = fmz[cm];
    double tvf = first ? 0.5 * (tv_km1 + tv_k) : tetavf[c];
    const double wk = w[c], fzf = fmzf[c];
    tvf = tvf - wk * fzf * dtrdz * (tv_km1 - tv_k);
    tetavf[c] = tvf;
    const double zrom1w = kCpd * tvf * fzf;
    double zwexpl = wk - zrom1w * dtrdz * (pai_km1 - pai_k) - kEgrav * dts;
    zwexpl = zwexpl + kRdrcv * zrom1w * dtrdz * (pai_km1 * z_km1 - pai_k * z_k);
    const double zu = zcs2 * fmz_km1 * zrom1w * pai_km1 + ffilt[k - 1];
    const double zd = zcs2 * fmz_k * zrom1w * pai_k + ffilt[k - 1];
    const double zrapp = 1.0 / (1.0 + zd + zu - zd * wk_up);
    const double wt = zrapp * (zwexpl + zd * w_up);
    const double wkk = zrapp * zu;
    wtb[c] = wt;
    wkb[c] = wkk;
    w_up = wt;
    wk_up = wkk;
    s_k = s_km1;
    z_k = z_km1;
    tv_k = tv_km1;
    pai_k = pai_km1;
    fmz_k = fmz_km1;
  }
  double below = w1;
  for (int k = 2; k <= kz; ++k) {
    const long long c = c1 + (long long)(k - 1) * P, cm = c - P;
    const double wf = wtb[c] + wkb[c] * below;
    w[c] = wf;
    pai[cm] = pai[cm] * (1.0 - kRdrcv * (z2b[cm] + (dtrdz * fmz[cm] * (below - wf))));
    below = wf;
  }
  pai[ckz] = pai[ckz] * (1.0 - kRdrcv * (z2b[ckz] + (dtrdz * fmz[ckz] * (below - wtop))));
}

// ---------------------------------------------------------------------------
// Acoustic sub-step, part 3: divergence damping (:738-765) and the
// horizontal momentum equations (:677-721).  The damping increment is added
// before the momentum increment exactly as the in-place reference sequence
// does; ud/vd are the sub-step's starting winds, read before either update.
// Each thread owns a chunk of levels of one column; 2-D metrics load once.
// The Exner differences are taken in double before entering the bulk type.
// The momentum update of one u point and one v point (:673-724 with the
// divergence damping :738-765), in double in both forms: the terrain-slope
// gravity term and the pressure gradient nearly cancel over slopes, and the
// difference needs more than float.  Operands in the reference's order.
struct MomK {
  double dts, dtrdx, dtrdy, dxrdt, zfz, hcp;
  int divdamp, rotllr;
};
// u at (j,i): pw, tw, zdw are pai, tetav, zdiv2 of (j-1,i); vo is v(j,i)
__device__ __forceinline__ double mom_u(const MomK &q, double uo, double vo, double pc,
                                        double pw, double tc, double tw, double zdc,
                                        double zdw, double bu, double xk, double gz,
                                        double muv, double hxv, double coruv) {
  double un = uo;
  if (q.divdamp) {
    const double xdam = q.dxrdt * xk * muv;
    un = un + xdam * (zdc - zdw);
  }
  const double zcx = q.dtrdx * muv;
  const double zrom1u = q.hcp * (tw + tc);
  const double zcor1u = coruv * q.dts * vo;
  return un + bu * (zcor1u - q.zfz * hxv * gz - zcx * zrom1u * (pc - pw));
}
// v at (j,i): ps, ts, zds are those of (j,i-1); uo is u(j,i)
__device__ __forceinline__ double mom_v(const MomK &q, double vo, double uo, double pc,
                                        double ps, double tc, double ts, double zdc,
                                        double zds, double bv, double xk, double gz,
                                        double mvv, double hyv, double corvv) {
  double vn = vo;
  if (q.divdamp) {
    const double xdam = q.rotllr ? q.dxrdt * xk : q.dxrdt * xk * mvv;
    vn = vn + xdam * (zdc - zds);
  }
  const double zcy = q.rotllr ? q.dtrdy : q.dtrdy * mvv;
  const double zrom1v = q.hcp * (ts + tc);
  const double zcor1v = corvv * q.dts * uo;
  return vn + bv * (-zcor1v - q.zfz * hyv * gz - zcy * zrom1v * (pc - ps));
}

template <typename R>
struct SoundMomentum {
  Dims d;
  R * u;
  R * v;
  const R * zdiv2;
  const double * tetav;
  const double * pai;
  BdyWt<R> bwu;
  BdyWt<R> bwv;
  const R * mu;
  const R * mv;
  const R * hx;
  const R * hy;
  const R * coru;
  const R * corv;
  const double * xkdamp;
  const double * gzitakh;
  double dts_;
  double dtrdx_;
  double dtrdy_;
  double dxrdt_;
  int divdamp;
  int rotllr;
  int kc;
  static __device__ __forceinline__ void run(const SoundMomentum &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SoundMomentum<R>::run(const SoundMomentum &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  R *__restrict__ u = args_.u;
  R *__restrict__ v = args_.v;
  const R *__restrict__ zdiv2 = args_.zdiv2;
  const double *__restrict__ tetav = args_.tetav;
  const double *__restrict__ pai = args_.pai;
  const BdyWt<R> bwu = args_.bwu;
  const BdyWt<R> bwv = args_.bwv;
  const R *__restrict__ mu = args_.mu;
  const R *__restrict__ mv = args_.mv;
  const R *__restrict__ hx = args_.hx;
  const R *__restrict__ hy = args_.hy;
  const R *__restrict__ coru = args_.coru;
  const R *__restrict__ corv = args_.corv;
  const double *__restrict__ xkdamp = args_.xkdamp;
  const double *__restrict__ gzitakh = args_.gzitakh;
  const double dts_ = args_.dts_;
  const double dtrdx_ = args_.dtrdx_;
  const double dtrdy_ = args_.dtrdy_;
  const double dxrdt_ = args_.dxrdt_;
  const int divdamp = args_.divdamp;
  const int rotllr = args_.rotllr;
  const int kc = args_.kc;

  const int j = d.jc1 + bx * TJ + threadIdx.x;
  const int i = d.ic1 + by * TI + threadIdx.y;
  const int k0 = 1 + (int)bz * kc;
  const int k1 = k0 + kc - 1 < d.kz ? k0 + kc - 1 : d.kz;
  if (j > d.je2 || i > d.ie2) return;
  // only the loads and stores are of the bulk type (mom_u, mom_v)
  const MomK q{dts_, dtrdx_, dtrdy_, dxrdt_, kEgrav * dts_, 0.5 * kCpd, divdamp, rotllr};
  const long long c2 = d.at2(j, i);
  const long long ow = d.oj(j, -1), os = d.oi(i, -1);   // west / south neighbours
  const bool du = i <= d.ic2;               // u point: jdi x ici
  const bool dv = j <= d.jc2;               // v point: jci x idi
  const double muv = mu[c2], mvv = mv[c2];
  const R hxr = hx[c2], hyr = hy[c2], corur = coru[c2], corvr = corv[c2];
  const R *__restrict__ bcu = bwu.col(c2);
  const R *__restrict__ bcv = bwv.col(c2);
  const double hxv = du ? (double)hxr : 0.0, hyv = dv ? (double)hyr : 0.0;
  const double coruv = du ? (double)corur : 0.0, corvv = dv ? (double)corvr : 0.0;
  // every load of a level is unconditional (the points outside jdi x ici or
  // jci x idi still read valid neighbours) and so is every store: a point
  // outside the u (v) range gets back the value it held (R -> double -> R
  // is exact), which no other thread of the pass reads.  A predicated store
  // would let the compiler sink its operands' loads into the branch, behind
  // the other store, and a level would take three memory round trips
  // instead of one.
  for (int k = k0; k <= k1; ++k) {
    const long long c = d.at(j, i, k);
    const R ur = u[c], vr = v[c];
    const double pc = pai[c], tc = tetav[c];
    const double pw = pai[c + ow], tw = tetav[c + ow];
    const double ps = pai[c + os], ts = tetav[c + os];
    const R zdcr = zdiv2[c], zdwr = zdiv2[c + ow], zdsr = zdiv2[c + os];
    const R bur = bcu[k - 1], bvr = bcv[k - 1];
    const double xk = xkdamp[k - 1], gz = gzitakh[k - 1];
    // every load of the level is issued before any use: left to itself the
    // scheduler places some after the first wait (a second round trip)
    __builtin_amdgcn_sched_barrier(0);
    const double uo = ur, vo = vr;
    const double zdc = zdcr, zdw = zdwr, zds = zdsr;
    const double bu = bur, bv = bvr;
    const double un = mom_u(q, uo, vo, pc, pw, tc, tw, zdc, zdw, bu, xk, gz, muv, hxv, coruv);
    const double vn = mom_v(q, vo, uo, pc, ps, tc, ts, zdc, zds, bv, xk, gz, mvv, hyv, corvv);
    u[c] = du ? R(un) : ur;
    v[c] = dv ? R(vn) : vr;
  }
}

// ---------------------------------------------------------------------------
// Advection prologue: complete s (:728-734), u,v -> ux,vx (:1524-1569) and
// w -> wx (:1445-1459) on every cross point.
template <typename R>
struct StaggerIn {
  Dims d;
  const R * u;
  const R * v;
  const double * w;
  const R * fmzf;
  R * s;
  R * ux;
  R * vx;
  R * wx;
  int uv_only;       // 1: only u,v -> ux,vx (uvstagtouvx of boundary())
  ZMet zm;
  int kc;            // levels per thread: levels 1 + bz*kc .. of the column
  static __device__ __forceinline__ void run(const StaggerIn &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void StaggerIn<R>::run(const StaggerIn &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const R *__restrict__ u = args_.u;
  const R *__restrict__ v = args_.v;
  const double *__restrict__ w = args_.w;
  const R *__restrict__ fmzf = args_.fmzf;
  R *__restrict__ s = args_.s;
  R *__restrict__ ux = args_.ux;
  R *__restrict__ vx = args_.vx;
  R *__restrict__ wx = args_.wx;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  const int k0 = 1 + bz * args_.kc;
  const int k1 = k0 + args_.kc - 1 < kz ? k0 + args_.kc - 1 : kz;
  if (j > d.je2 || i > d.ie2) return;
  const long long pj = d.pj, P = d.plane;
  const R h = R(0.5), a9 = R(0.5625), a1 = R(0.0625);
  // every load unconditional (neighbour offsets clamped into the arrays, or
  // wrapped on a periodic domain); the edge formulas of has_bdy* pick from
  // the same loaded values (a + b == b + a)
  long long ou1, ou2, oum, ov1, ov2, ovm;
  bool jedge, iedge;
  if (d.crm) {
    ou1 = d.oj(j, 1); ou2 = d.oj(j, 2); oum = d.oj(j, -1);
    ov1 = d.oi(i, 1); ov2 = d.oi(i, 2); ovm = d.oi(i, -1);
    jedge = iedge = false;
  } else {
    ou1 = 1; ou2 = j + 2 <= jx ? 2 : 1; oum = j >= 2 ? -1 : 0;
    ov1 = pj; ov2 = i + 2 <= iy ? 2 * pj : pj; ovm = i >= 2 ? -pj : 0;
    jedge = j == 1 || j == jx - 1;
    iedge = i == 1 || i == iy - 1;
  }
  const ZMet zm = args_.zm;
  const bool uvw = !args_.uv_only;
  const bool inner = j >= d.jc1 && j <= d.jc2 && i >= d.ic1 && i <= d.ic2;
  const double zh = (uvw && zm.on) ? zm.hc[d.at2(j, i)] : 0.0;
  for (int k = k0; k <= k1; ++k) {
    const long long c = d.at(j, i, k);
    const R u0 = u[c], u1 = u[c + ou1], u2 = u[c + ou2], um = u[c + oum];
    const R v0 = v[c], v1 = v[c + ov1], v2 = v[c + ov2], vm = v[c + ovm];
    // the loads of w, s and fmzf (or its orography operands) are issued with
    // those of u and v and waited for together: behind the ux, vx stores
    // they would wait for those stores too (one counter covers loads and
    // stores)
    double w0d = 0.0, w1d = 0.0, w2d = 0.0, wmd = 0.0, zg = 0.0, zb = 0.0;
    R sc = R(0), fzs = R(0);
    if (uvw) {
      const long long cw2 = k + 2 <= kz + 1 ? c + 2 * P : c + P, cwm = k >= 2 ? c - P : c;
      w0d = w[c]; w1d = w[c + P]; w2d = w[cw2]; wmd = w[cwm];
      sc = s[c];
      if (zm.on) {
        // ZMet::fmzf(c2, k) = 1 / den(ht, c2, kz + k - 1)
        zg = zm.ga[kz + k - 1];
        zb = zm.ba[kz + k - 1];
      } else {
        fzs = fmzf[c];
      }
    }
    issued(u0, u1, u2, um);
    issued(v0, v1, v2, vm);
    issued(w0d, w1d, w2d, wmd);
    issued(sc, fzs, zh, zg, zb);
    ux[c] = jedge ? h * (u0 + u1) : a9 * (u1 + u0) - a1 * (u2 + um);
    vx[c] = iedge ? h * (v0 + v1) : a9 * (v1 + v0) - a1 * (v2 + vm);
    if (!uvw) continue;
    const R w0 = R(w0d), w1 = R(w1d), w2 = R(w2d), wm = R(wmd);
    wx[c] = (k == 1 || k == kz) ? h * (w1 + w0) : a9 * (w1 + w0) - a1 * (w2 + wm);
    if (inner) {
      const R fz = zm.on ? R(1.0 / (zg * (zh * kRegrav) + zb)) : fzs;
      s[c] = k >= 2 ? (w0 + sc) * fz : R(0);
      if (k == kz) s[c + P] = R(0);
    }
  }
}

// ---------------------------------------------------------------------------
// WAF vertical advection, both half steps (:862-922), one column and one
// field per lane, streamed top to bottom through register windows: the
// first half step emits wz1(k) once pp(k+2) is loaded, the second emits
// wz2(k-2) once wz1(k) is known, so every face flux is evaluated once and
// each input is read once.  zrfmu(k) = dtrdz*fmz(k)/fmzf(k) and
// zrfmd(k) = dtrdz*fmz(k)/fmzf(k+1) are precomputed at init with exactly
// these operations.  T is the field/arithmetic type, S the type of s and
// of the metric ratios.
template <typename T>
__device__ __forceinline__ T waf_vflux(T skp1, T dtrdz, T xm1, T x0, T x1,
                                       T x2) {
  // face kk+1 with xm1 = x(max(kk-1,1)), x0 = x(kk), x1 = x(kk+1),
  // x2 = x(min(kk+2,kz)): the k1/k1p1 clamps of :870-880
  const T zamu = skp1 * dtrdz;
  T is, num;
  if (zamu >= T(0)) {
    is = T(1);
    num = x1 - x2;
  } else {
    is = T(-1);
    num = xm1 - x0;
  }
  const T r = flow_param(num, x0 - x1);
  const T b = waf_b(r);
  const T zphi = is + zamu * b - is * b;
  return T(0.5) * skp1 * ((T(1) + zphi) * x1 + (T(1) - zphi) * x0);
}

template <typename T, typename S, int G>
struct WafVertical {
  Dims d;
  AdvPtrs<T> ap;
  const S * s;
  const S * zrfmu;
  const S * zrfmd;
  double dtrdz_;
  int nf;                      // fields in the batch; a thread takes G of them
  static __device__ __forceinline__ void run(const WafVertical &args_, int bx, int by, int bz);
};

// A thread streams the column of G fields (f0 .. f0+G-1, f0 = G*bz) so that
// the per-level loads of s, zrfmu and zrfmd are shared by the group.
template <typename T, typename S, int G>
__device__ __forceinline__ void WafVertical<T, S, G>::run(const WafVertical &args_, int bx, int by,
                                                          int bz) {
  const Dims d = args_.d;
  const AdvPtrs<T> &ap = args_.ap;
  const S *__restrict__ s = args_.s;
  const S *__restrict__ zrfmu = args_.zrfmu;
  const S *__restrict__ zrfmd = args_.zrfmd;
  const double dtrdz_ = args_.dtrdz_;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int f0 = (int)bz * G;
  const int ng = args_.nf - f0 < G ? args_.nf - f0 : G;
  if (j > d.je2 || i > d.ie2) return;
  const T dtrdz = T(dtrdz_);
  const int kz = d.kz;
  const long long P = d.plane;
  const long long c1 = d.at(j, i, 1);
  const S *__restrict__ ps = s + c1;
  const S *__restrict__ pu = zrfmu + c1;
  const S *__restrict__ pd = zrfmd + c1;
  const T *px[G];
  T *po[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int f = f0 + (g < ng ? g : 0);
    px[g] = ap.src[f] + c1;
    po[g] = ap.dst[f] + c1;
  }
  // per field: first half step window x(max(k-1,1)), x(k), x(k+1),
  // x(min(k+2,kz)); wfw1 on the top face of level k; wz1 history; wfw2
  T a_m1[G], a_0[G], a_1[G], a_2[G], f1[G], h1[G], h2[G], h3[G], f2[G];
  const long long k3 = (long long)((3 <= kz ? 3 : kz) - 1) * P;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g < ng) {
      a_m1[g] = px[g][0];
      a_0[g] = a_m1[g];
      a_1[g] = px[g][P];
      a_2[g] = px[g][k3];
    }
    f1[g] = T(0);
    h1[g] = h2[g] = h3[g] = T(0);
    f2[g] = T(0);
  }
  T s_k = T(ps[0]);
  T sm1 = T(0), sm2 = T(0);              // s(k-1), s(k-2)
  T um1 = T(0), um2 = T(0), dm1 = T(0), dm2 = T(0);
  // s(k+1), zrfmu(k), zrfmd(k) of the next level are requested a level
  // ahead, before this level's stores (a load issued after a store would
  // wait for the store as well)
  S nx_s = ps[P], nx_u = pu[0], nx_d = pd[0];
  for (int k = 1; k <= kz; ++k) {
    const long long o = (long long)(k - 1) * P;
    const T s_kp1 = T(nx_s);
    const T zu = T(nx_u), zd = T(nx_d);
    {
      const long long on = k < kz ? o + P : o;   // level k + 1, clamped
      nx_s = ps[on + P];
      nx_u = pu[on];
      nx_d = pd[on];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (g < ng) {
        const T f1n = (k < kz) ? waf_vflux(s_kp1, dtrdz, a_m1[g], a_0[g], a_1[g], a_2[g]) : T(0);
        const T zdv = (s_k * zu - s_kp1 * zd) * a_0[g];
        const T w1 = a_0[g] - f1[g] * zu + f1n * zd + zdv;     // wz1(k)
        f1[g] = f1n;
        {
          const T t = a_2[g];
          a_m1[g] = a_0[g];
          a_0[g] = a_1[g];
          a_1[g] = t;
          a_2[g] = (k + 3 <= kz) ? px[g][(long long)(k + 2) * P] : t;
        }
        if (k >= 3) {                      // emit wz2(k-2)
          const int kp = k - 2;
          const T ym1 = (kp == 1) ? h2[g] : h3[g];
          const T y0 = h2[g], y1 = h1[g], y2 = w1;
          const T f2n = waf_vflux(sm1, dtrdz, ym1, y0, y1, y2);
          const T zdv2 = (sm2 * um2 - sm1 * dm2) * y0;
          po[g][(long long)(kp - 1) * P] = y0 - f2[g] * um2 + f2n * dm2 + zdv2;
          f2[g] = f2n;
        }
        h3[g] = h2[g];
        h2[g] = h1[g];
        h1[g] = w1;
      }
    }
    sm2 = sm1;
    sm1 = s_k;
    s_k = s_kp1;
    um2 = um1;
    um1 = zu;
    dm2 = dm1;
    dm1 = zd;
  }
  // flush wz2(kz-1) and wz2(kz): wz1 history is (h3, h2, h1) =
  // (wz1(kz-2), wz1(kz-1), wz1(kz)); metrics of kz-1 in (sm2, um2, dm2), of
  // kz in (sm1, um1, dm1); s(kz+1) in s_k
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g < ng) {
      {
        const int kp = kz - 1;
        const T ym1 = (kp == 1) ? h2[g] : h3[g];
        const T y0 = h2[g], y1 = h1[g], y2 = h1[g];
        const T f2n = waf_vflux(sm1, dtrdz, ym1, y0, y1, y2);   // s(kz)
        const T zdv2 = (sm2 * um2 - sm1 * dm2) * y0;
        po[g][(long long)(kp - 1) * P] = y0 - f2[g] * um2 + f2n * dm2 + zdv2;
        f2[g] = f2n;
      }
      {
        const T y0 = h1[g];                  // wz1(kz)
        const T f2n = T(0);                  // wfw2(kz+1)
        const T zdv2 = (sm1 * um1 - s_k * dm1) * y0;
        po[g][(long long)(kz - 1) * P] = y0 - f2[g] * um1 + f2n * dm1 + zdv2;
      }
    }
  }
}

// zrfmu = dtrdz*fmz(k)/fmzf(k), zrfmd = dtrdz*fmz(k)/fmzf(k+1) (:888-889),
// from the double metrics, stored in the bulk type.
template <typename R>
struct WafMetrics {
  Dims d;
  const double * fmz;
  const double * fmzf;
  R * zrfmu;
  R * zrfmd;
  double dtrdz;
  static __device__ __forceinline__ void run(const WafMetrics &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void WafMetrics<R>::run(const WafMetrics &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const double *__restrict__ fmz = args_.fmz;
  const double *__restrict__ fmzf = args_.fmzf;
  R *__restrict__ zrfmu = args_.zrfmu;
  R *__restrict__ zrfmd = args_.zrfmd;
  const double dtrdz = args_.dtrdz;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int k = 1 + bz;
  if (j > d.je2 || i > d.ie2) return;
  const long long c = d.at(j, i, k);
  zrfmu[c] = R(dtrdz * fmz[c] / fmzf[c]);
  zrfmd[c] = R(dtrdz * fmz[c] / fmzf[c + d.plane]);
}

// ---------------------------------------------------------------------------
// WAF meridional + zonal passes (:926-1038) of a batch of advected fields at
// one level on a 64x4 tile, fused through LDS: wz (halo 2 in i and j) ->
// zpby -> p0 (halo 2 in j) -> zpbw -> new field.  The field-independent
// parts of each point (Courant number, upwind row/column, 0.5*v, 0.5*u,
// metric products and the divergence-correction coefficient, each formed
// with the reference's operand order in the winds' type S) are computed
// once per tile in registers and reused for all fields of the batch.  A
// field is read from src and written to dst (the halo points of p0 need the
// old field at points a neighbouring tile owns); boundary points are copied
// unchanged.  T is the field/arithmetic type, S the type of the winds and
// metrics (T = S = double in the FP64 form).
constexpr int kThr = TJ * TI;                // threads of a block (64 x 4)

// PER = 1 for a doubly periodic domain (halo loads wrap); its own
// instantiation, so that the bounded form keeps constant offsets.
template <typename T, typename S, int TR, int PER>
struct WafHorizontal {
  Dims d;
  AdvPtrs<T> ap;
  AdvPtrs<T> wzp;
  int nf;
  int ngroups;                 // field groups: blockIdx.x = tile column * ngroups + group
  int fgroup;                  // fields per group
  const S * u;
  const S * v;
  const S * mu;
  const S * mv;
  const S * mx;
  const S * mx2;
  const S * rmu;
  const S * rmv;
  const S * rfmzu;
  const S * rfmzv;
  const S * fmz;
  double dtrdx_;
  double dtrdy_;
  int rotllr;
  ZMet zm;
  // the curvature terms of advection (:811-825) fused into the output of
  // the batch's ux (field cux) and vx (field cvx, after cux in the same
  // block); -1 when the curvature pass runs on its own
  int cux, cvx;
  const S * tanrot;
  double rdx_, dta_;
  static __device__ __forceinline__ void run(const WafHorizontal &args_, int bx, int by, int bz);
};

// The body of a horizontal-WAF tile.  The field-independent coefficients
// (Courant numbers, upwind rows / columns, 0.5 v, 0.5 u, metric products,
// divergence-correction coefficients) are formed once per tile in the winds'
// type S, with the reference's operand order, and each field of type F uses
// F(coefficient).  nf fields of type T are advected, then nf2 fields of type
// T2 (the mixed form's double batch in the fused launch; nf2 = 0 otherwise).
// The LDS tiles are sized for T2 and hold either type.
template <typename T, typename T2, typename S, int TR, int PER>
__device__ __forceinline__ void waf_h_body(const WafHorizontal<T, S, TR, PER> &args_,
                                           const AdvPtrs<T2> *ap2, const AdvPtrs<T2> *wzp2,
                                           int nf2, int bx, int by, int bz) {
  // a tile of TJ x TR output points, TR / TI rows per thread
  constexpr int kWzN = (TR + 4) * (TJ + 4);    // wz tile
  constexpr int kPyN = (TR + 1) * (TJ + 4);    // zpby points
  constexpr int kP0N = TR * (TJ + 4);          // p0 points
  constexpr int kPwN = TR * (TJ + 1);          // zpbw points
  constexpr int kPyM = (kPyN + kThr - 1) / kThr;
  constexpr int kP0M = (kP0N + kThr - 1) / kThr;
  constexpr int kPwM = (kPwN + kThr - 1) / kThr;
  constexpr int kOM = TR / TI;
  const Dims d = args_.d;
  // fields [f0, f0 + nfg) of the batch; the groups of one tile are adjacent
  // blocks, so they share the tile's winds and metrics in L2
  const int grp = bx % args_.ngroups;
  bx = bx / args_.ngroups;
  const int f0 = grp * args_.fgroup;
  const int nf = (args_.nf - f0 < args_.fgroup ? args_.nf - f0 : args_.fgroup) + f0;
  const S *__restrict__ u = args_.u;
  const S *__restrict__ v = args_.v;
  const S *__restrict__ mu = args_.mu;
  const S *__restrict__ mv = args_.mv;
  const S *__restrict__ mx = args_.mx;
  const S *__restrict__ mx2 = args_.mx2;
  const S *__restrict__ rmu = args_.rmu;
  const S *__restrict__ rmv = args_.rmv;
  const S *__restrict__ rfmzu = args_.rfmzu;
  const S *__restrict__ rfmzv = args_.rfmzv;
  const S *__restrict__ fmz = args_.fmz;
  const int rotllr = args_.rotllr;
  // fmz, rfmzu, rfmzv at this level: formed from the orography when the
  // handle verified that, else streamed
  const ZMet zm = args_.zm;

  __shared__ T2 lwz[TR + 4][TJ + 4];
  __shared__ T2 lpp[TR][TJ + 4];
  __shared__ T2 lzpby[TR + 1][TJ + 4];
  __shared__ T2 lp0[TR][TJ + 4];
  __shared__ T2 lzpbw[TR][TJ + 1];
  const S dtrdx = S(args_.dtrdx_), dtrdy = S(args_.dtrdy_);
  const int k = 1 + (int)bz;
  const int j0 = 1 + bx * TJ, i0 = 1 + by * TR;
  const int tid = threadIdx.y * TJ + threadIdx.x;
  constexpr int crm = PER;
  const int jce2 = PER ? d.je2 : d.jx - 1, ice2 = PER ? d.ie2 : d.iy - 1;
  // interior extents (jc1..jc2, ic1..ic2) of this form
  const int jc1 = PER ? 1 : 2, ic1 = PER ? 1 : 2;
  const int jc2 = PER ? jce2 : jce2 - 1, ic2 = PER ? ice2 : ice2 - 1;
  // upwind clamps (init_moloch :280-293): two points wider on a periodic
  // domain, where they never bind
  const int jmin = crm ? -1 : 1, jmax = crm ? jce2 + 2 : jce2;
  const int imin = crm ? -1 : 1, imax = crm ? ice2 + 2 : ice2;
  // the j range of the tile's halo columns (wz, pp, zpby, p0) and the i range
  // of its halo rows: the cross extent, or two more each side when periodic
  const int hj1 = crm ? -1 : 1, hj2 = crm ? jce2 + 2 : jce2;
  const int hi1 = crm ? -1 : 1, hi2 = crm ? ice2 + 2 : ice2;

  // ---- field-independent coefficients --------------------------------
  // Every operand of the tile's coefficients is loaded first, in straight-
  // line code: a point outside its range reads element (1,1,1) or a small
  // in-range offset from it, and its coefficients are zeroed afterwards.
  // The loads are then in flight together; loaded inside per-point branches
  // (and behind the branch on each Courant number's sign), each point's
  // loads waited for the previous point's, some fifty round trips per tile.
  // The meridional fluxes take rfmzv with ROTLLR, else rfmzu (:1004-1005);
  // mx with ROTLLR, else mx2.
  // With double winds a formed metric is loaded as its orography value and
  // formed (as ZMet: fmz = 1/den, rfmzu = rfmzv = 1/(1/den)) after all loads
  // are issued; with float winds, where double raw values would cost two
  // waves of occupancy, it is formed as it is loaded.
  const S *__restrict__ rfmy = rotllr ? rfmzv : rfmzu;
  const double *hmy = rotllr ? zm.hv : zm.hu;
  const S *__restrict__ mxs = rotllr ? mx : mx2;
  constexpr bool kRaw = sizeof(S) == sizeof(double);
  using M = typename std::conditional<kRaw, double, S>::type;
  auto l_fz = [&](long long c, long long c2) -> M {
    if constexpr (kRaw) return zm.on ? zm.hc[c2] : M(fmz[c]);
    else return zm.on ? S(zm.fmz(c2, k)) : fmz[c];
  };
  auto l_ry = [&](long long c, long long c2) -> M {
    if constexpr (kRaw) return zm.on ? hmy[c2] : M(rfmy[c]);
    else return zm.on ? S(zm.rfm(hmy, c2, k)) : rfmy[c];
  };
  auto l_ru = [&](long long c, long long c2) -> M {
    if constexpr (kRaw) return zm.on ? zm.hu[c2] : M(rfmzu[c]);
    else return zm.on ? S(zm.rfmzu(c2, k)) : rfmzu[c];
  };
  const double zga = zm.on ? zm.ga[k - 1] : 0.0, zba = zm.on ? zm.ba[k - 1] : 0.0;
  auto f_fz = [&](M raw) -> S {
    if constexpr (kRaw) return zm.on ? S(1.0 / (zga * (raw * kRegrav) + zba)) : S(raw);
    else return raw;
  };
  auto f_rf = [&](M raw) -> S {
    if constexpr (kRaw) return zm.on ? S(1.0 / (1.0 / (zga * (raw * kRegrav) + zba))) : S(raw);
    else return raw;
  };
  S y_zamu[kPyM], y_hv[kPyM], y_is[kPyM];
  int y_rh[kPyM], y_rhm1[kPyM];
  bool y_ok[kPyM];
  S ly_v[kPyM], ly_mv[kPyM];
#pragma unroll
  for (int m = 0; m < kPyM; ++m) {
    const int n = tid + m * kThr;
    const int r = n / (TJ + 4), q = n % (TJ + 4);
    const int ii = i0 + r, jj = j0 - 2 + q;
    // zpby on i = ici1 .. ice2ga (one row past ice2 when periodic)
    y_ok[m] = n < kPyN && ii >= ic1 && ii <= ice2 + crm && jj >= hj1 && jj <= hj2;
    const long long c = y_ok[m] ? (PER ? d.atw(jj, ii, k) : d.at(jj, ii, k)) : 0;
    const long long c2 = y_ok[m] ? (PER ? d.at2w(jj, ii) : d.at2(jj, ii)) : 0;
    ly_v[m] = v[c];
    ly_mv[m] = mv[c2];
  }
  S p_zrfms[kP0M], p_zrfmn[kP0M], p_cdv[kP0M], p_mf[kP0M];
  bool p_ok[kP0M];
  M lp_fz[kP0M], lp_rn[kP0M], lp_rs[kP0M];
  S lp_vn[kP0M], lp_vs[kP0M], lp_mn[kP0M], lp_ms[kP0M], lp_mx[kP0M];
#pragma unroll
  for (int m = 0; m < kP0M; ++m) {
    const int n = tid + m * kThr;
    const int r = n / (TJ + 4), q = n % (TJ + 4);
    const int ii = i0 + r, jj = j0 - 2 + q;
    p_ok[m] = n < kP0N && ii >= ic1 && ii <= ic2 && jj >= hj1 && jj <= hj2;
    const long long c = p_ok[m] ? (PER ? d.atw(jj, ii, k) : d.at(jj, ii, k)) : 0;
    const long long c2 = p_ok[m] ? (PER ? d.at2w(jj, ii) : d.at2(jj, ii)) : 0;
    // the row north
    const long long pj = (PER && p_ok[m]) ? d.oi(d.wi(ii), 1) : (long long)d.pj;
    lp_fz[m] = l_fz(c, c2);
    lp_rn[m] = l_ry(c + pj, c2 + pj);
    lp_rs[m] = l_ry(c, c2);
    lp_vn[m] = v[c + pj];
    lp_vs[m] = v[c];
    lp_mn[m] = rmv[c2 + pj];
    lp_ms[m] = rmv[c2];
    lp_mx[m] = mxs[c2];
  }
  S w_zamu[kPwM], w_hu[kPwM], w_is[kPwM];
  int w_ch[kPwM], w_chm1[kPwM];
  bool w_ok[kPwM];
  S lw_u[kPwM], lw_mu[kPwM];
#pragma unroll
  for (int m = 0; m < kPwM; ++m) {
    const int n = tid + m * kThr;
    const int r = n / (TJ + 1), q = n % (TJ + 1);
    const int ii = i0 + r, jj = j0 + q;
    // zpbw on j = jci1 .. jce2ga (one column past jce2 when periodic)
    w_ok[m] = n < kPwN && ii >= ic1 && ii <= ic2 && jj >= jc1 && jj <= jce2 + crm;
    const long long c = w_ok[m] ? (PER ? d.atw(jj, ii, k) : d.at(jj, ii, k)) : 0;
    const long long c2 = w_ok[m] ? (PER ? d.at2w(jj, ii) : d.at2(jj, ii)) : 0;
    lw_u[m] = u[c];
    lw_mu[m] = mu[c2];
  }
  S o_zrfmw[kOM], o_zrfme[kOM], o_cdv[kOM], o_mf[kOM];
  bool o_in[kOM], o_int[kOM];
  long long oc[kOM];
  M lo_fz[kOM], lo_rw[kOM], lo_re[kOM];
  S lo_ue[kOM], lo_uw[kOM], lo_me[kOM], lo_mw[kOM], lo_mx[kOM];
#pragma unroll
  for (int mo = 0; mo < kOM; ++mo) {
    const int j = j0 + threadIdx.x, i = i0 + threadIdx.y + TI * mo;
    o_in[mo] = j <= jce2 && i <= ice2;
    o_int[mo] = j >= jc1 && j <= jc2 && i >= ic1 && i <= ic2;
    oc[mo] = o_in[mo] ? d.at(j, i, k) : 0;
    const long long c = o_int[mo] ? oc[mo] : 0;
    const long long c2 = o_int[mo] ? d.at2(j, i) : 0;
    const long long e = (PER && o_int[mo]) ? d.oj(j, 1) : 1;   // east
    lo_fz[mo] = l_fz(c, c2);
    lo_rw[mo] = l_ru(c, c2);
    lo_re[mo] = l_ru(c + e, c2 + e);
    lo_ue[mo] = u[c + e];
    lo_uw[mo] = u[c];
    lo_me[mo] = rmu[c2 + e];
    lo_mw[mo] = rmu[c2];
    lo_mx[mo] = mxs[c2];
  }
  // the coefficients, with the reference's operand order
#pragma unroll
  for (int m = 0; m < kPyM; ++m) {
    const int n = tid + m * kThr;
    const int ii = i0 + n / (TJ + 4);
    const S vv = ly_v[m];
    const S zamu = rotllr ? vv * dtrdy : vv * ly_mv[m] * dtrdy;
    const bool pos = zamu > S(0);
    const int ih = pos ? ii - 1 : (ii + 1 < imax ? ii + 1 : imax);
    const int ihm1 = ih - 1 > imin ? ih - 1 : imin;
    y_zamu[m] = y_ok[m] ? zamu : S(0);
    y_hv[m] = y_ok[m] ? S(0.5) * vv : S(0);
    y_is[m] = y_ok[m] ? (pos ? S(1) : S(-1)) : S(0);
    y_rh[m] = y_ok[m] ? ih - (i0 - 2) : 0;
    y_rhm1[m] = y_ok[m] ? ihm1 - (i0 - 2) : 0;
  }
#pragma unroll
  for (int m = 0; m < kP0M; ++m) {
    const S fz = f_fz(lp_fz[m]), rn = f_rf(lp_rn[m]), rs = f_rf(lp_rs[m]);
    S zrfmn, zrfms, cdv, mf = S(1);
    if (rotllr) {
      const S zhxvtn = dtrdy * lp_mn[m] * lp_mx[m];
      const S zhxvts = dtrdy * lp_ms[m] * lp_mx[m];
      zrfmn = zhxvtn * fz * rn;
      zrfms = zhxvts * fz * rs;
      cdv = lp_vn[m] * zrfmn - lp_vs[m] * zrfms;
    } else {
      zrfmn = dtrdy * fz * rn;
      zrfms = dtrdy * fz * rs;
      cdv = lp_vn[m] * lp_mn[m] * zrfmn - lp_vs[m] * lp_ms[m] * zrfms;
      mf = lp_mx[m];
    }
    p_zrfmn[m] = p_ok[m] ? zrfmn : S(0);
    p_zrfms[m] = p_ok[m] ? zrfms : S(0);
    p_cdv[m] = p_ok[m] ? cdv : S(0);
    p_mf[m] = p_ok[m] ? mf : S(1);
  }
#pragma unroll
  for (int m = 0; m < kPwM; ++m) {
    const int n = tid + m * kThr;
    const int jj = j0 + n % (TJ + 1);
    const S uu = lw_u[m];
    const S zamu = uu * lw_mu[m] * dtrdx;
    const bool pos = zamu > S(0);
    const int jh = pos ? jj - 1 : (jj + 1 < jmax ? jj + 1 : jmax);
    const int jhm1 = jh - 1 > jmin ? jh - 1 : jmin;
    w_zamu[m] = w_ok[m] ? zamu : S(0);
    w_hu[m] = w_ok[m] ? S(0.5) * uu : S(0);
    w_is[m] = w_ok[m] ? (pos ? S(1) : S(-1)) : S(0);
    w_ch[m] = w_ok[m] ? jh - (j0 - 2) : 0;
    w_chm1[m] = w_ok[m] ? jhm1 - (j0 - 2) : 0;
  }
#pragma unroll
  for (int mo = 0; mo < kOM; ++mo) {
    const S fz = f_fz(lo_fz[mo]), rw = f_rf(lo_rw[mo]), re = f_rf(lo_re[mo]);
    S zrfmw, zrfme, cdv, mf = S(1);
    if (rotllr) {
      const S zcostx = dtrdx * lo_mx[mo];
      zrfmw = zcostx * fz * rw;
      zrfme = zcostx * fz * re;
      cdv = lo_ue[mo] * zrfme - lo_uw[mo] * zrfmw;
    } else {
      zrfmw = dtrdx * fz * rw;
      zrfme = dtrdx * fz * re;
      cdv = lo_ue[mo] * lo_me[mo] * zrfme - lo_uw[mo] * lo_mw[mo] * zrfmw;
      mf = lo_mx[mo];
    }
    o_zrfmw[mo] = o_int[mo] ? zrfmw : S(0);
    o_zrfme[mo] = o_int[mo] ? zrfme : S(0);
    o_cdv[mo] = o_int[mo] ? cdv : S(0);
    o_mf[mo] = o_int[mo] ? mf : S(1);
  }

  // ---- per field ----------------------------------------------------------
  // The wz / pp tile of field f+1 is loaded into registers while field f is
  // processed, so the global-load latency overlaps the LDS stages.
  constexpr int kWzM = (kWzN + kThr - 1) / kThr;
  long long t_c[kWzM];
  bool t_ok[kWzM], t_pp[kWzM];
#pragma unroll
  for (int m = 0; m < kWzM; ++m) {
    const int n = tid + m * kThr;
    const int r = n / (TJ + 4), q = n % (TJ + 4);
    const int ii = i0 - 2 + r, jj = j0 - 2 + q;
    t_ok[m] = n < kWzN && ii >= hi1 && ii <= hi2 && jj >= hj1 && jj <= hj2;
    t_pp[m] = t_ok[m] && r >= 2 && r < TR + 2;
    t_c[m] = t_ok[m] ? (PER ? d.atw(jj, ii, k) : d.at(jj, ii, k)) : 0;
  }
  // fields fb .. fe - 1 of a batch of type F
  auto batch = [&](auto tag, const auto &ap, const auto &wzp, int fb, int fe) {
    using F = decltype(tag);
    F (*swz)[TJ + 4] = reinterpret_cast<F (*)[TJ + 4]>(&lwz[0][0]);
    F (*spp)[TJ + 4] = reinterpret_cast<F (*)[TJ + 4]>(&lpp[0][0]);
    F (*szpby)[TJ + 4] = reinterpret_cast<F (*)[TJ + 4]>(&lzpby[0][0]);
    F (*sp0)[TJ + 4] = reinterpret_cast<F (*)[TJ + 4]>(&lp0[0][0]);
    F (*szpbw)[TJ + 1] = reinterpret_cast<F (*)[TJ + 1]>(&lzpbw[0][0]);
    F nx_wz[kWzM], nx_pp[kWzM];
    auto prefetch = [&](int f) {
      const F *__restrict__ pp = ap.src[f];
      const F *__restrict__ wz = wzp.src[f];
#pragma unroll
      for (int m = 0; m < kWzM; ++m) {
        nx_wz[m] = t_ok[m] ? wz[t_c[m]] : F(0);
        nx_pp[m] = t_pp[m] ? pp[t_c[m]] : F(0);
      }
    };
    if (fe > fb) prefetch(fb);
    for (int f = fb; f < fe; ++f) {
#pragma unroll
      for (int m = 0; m < kWzM; ++m) {
        if (t_ok[m]) {
          const int n = tid + m * kThr;
          const int r = n / (TJ + 4), q = n % (TJ + 4);
          swz[r][q] = nx_wz[m];
          if (t_pp[m]) spp[r - 2][q] = nx_pp[m];
        }
      }
      if (f + 1 < fe) prefetch(f + 1);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kPyM; ++m) {
        if (y_ok[m]) {
          const int n = tid + m * kThr;
          const int r = n / (TJ + 4), q = n % (TJ + 4);
          const F wi = swz[r + 2][q], wim1 = swz[r + 1][q];
          const F rr = flow_param(swz[y_rh[m]][q] - swz[y_rhm1[m]][q], wi - wim1);
          const F b = waf_b(rr);
          const F is = F(y_is[m]);
          const F zphi = is + F(y_zamu[m]) * b - is * b;
          szpby[r][q] = F(y_hv[m]) * ((F(1) + zphi) * wim1 + (F(1) - zphi) * wi);
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kP0M; ++m) {
        if (p_ok[m]) {
          const int n = tid + m * kThr;
          const int r = n / (TJ + 4), q = n % (TJ + 4);
          const F zdv = F(p_cdv[m]) * spp[r][q];
          const F x = szpby[r][q] * F(p_zrfms[m]) - szpby[r + 1][q] * F(p_zrfmn[m]) + zdv;
          // ROTLLR has no mx2 factor; p_mf = 1 there and 1*x == x exactly
          sp0[r][q] = swz[r + 2][q] + F(p_mf[m]) * x;
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kPwM; ++m) {
        if (w_ok[m]) {
          const int n = tid + m * kThr;
          const int r = n / (TJ + 1), q = n % (TJ + 1);
          const F pjj = sp0[r][q + 2], pjm1 = sp0[r][q + 1];
          const F rr = flow_param(sp0[r][w_ch[m]] - sp0[r][w_chm1[m]], pjj - pjm1);
          const F b = waf_b(rr);
          const F is = F(w_is[m]);
          const F zphi = is + F(w_zamu[m]) * b - is * b;
          szpbw[r][q] = F(w_hu[m]) * ((F(1) + zphi) * pjm1 + (F(1) - zphi) * pjj);
        }
      }
      __syncthreads();
#pragma unroll
      for (int mo = 0; mo < kOM; ++mo) {
        if (o_in[mo]) {
          const int r = threadIdx.y + TI * mo;
          const F pold = spp[r][threadIdx.x + 2];
          F val = pold;
          if (o_int[mo]) {
            const F p0 = sp0[r][threadIdx.x + 2];
            const F zw = szpbw[r][threadIdx.x];
            const F ze = szpbw[r][threadIdx.x + 1];
            const F zdv = F(o_cdv[mo]) * pold;
            if (rotllr)
              val = p0 + zw * F(o_zrfmw[mo]) - ze * F(o_zrfme[mo]) + zdv;
            else
              val = p0 + F(o_mf[mo]) * (zw * F(o_zrfmw[mo]) - ze * F(o_zrfme[mo]) + zdv);
          }
          ap.dst[f][oc[mo]] = val;
        }
      }
      __syncthreads();
    }
  };
  batch(T(0), args_.ap, args_.wzp, f0, nf);
  // the curvature terms (Curvature, :811-825) on this block's output points
  // of ux and vx, which this thread has just stored; in the winds' type
  if (args_.cux >= f0 && args_.cux < nf) {
    T *pux = args_.ap.dst[args_.cux];
    T *pvx = args_.ap.dst[args_.cvx];
    const S rdx = S(args_.rdx_), dta = S(args_.dta_);
#pragma unroll
    for (int mo = 0; mo < kOM; ++mo) {
      if (o_int[mo]) {
        const int j = j0 + threadIdx.x, i = i0 + threadIdx.y + TI * mo;
        const long long c2 = d.at2(j, i);
        S tanx, tany;
        if (rotllr) {
          tanx = args_.tanrot[c2];
          tany = tanx;
        } else {
          tanx = (mu[c2 + d.oj(j, -1)] - mu[c2]) * rdx;
          tany = (mv[c2 + d.oi(i, -1)] - mv[c2]) * rdx;
        }
        T a = pux[oc[mo]];
        const T b = pvx[oc[mo]];
        a = a + a * b * T(tanx) * T(dta);
        pux[oc[mo]] = a;
        pvx[oc[mo]] = b - a * a * T(tany) * T(dta);
      }
    }
  }
  if constexpr (!std::is_same<T, T2>::value) batch(T2(0), *ap2, *wzp2, 0, nf2);
}

template <typename T, typename S, int TR, int PER>
__device__ __forceinline__ void WafHorizontal<T, S, TR, PER>::run(const WafHorizontal &args_, int bx,
                                                             int by, int bz) {
  waf_h_body<T, T, S, TR, PER>(args_, nullptr, nullptr, 0, bx, by, bz);
}

// The mixed form's double batch (pai, tetav) advected in the float batch's
// blocks, after the float fields, with the tile's one coefficient setup.
// Every block takes the whole batch (no field groups).
template <typename S, int TR, int PER>
struct WafHorizontalFused {
  WafHorizontal<float, S, TR, PER> f;
  AdvPtrs<double> ap2;
  AdvPtrs<double> wzp2;
  int nf2;
  static __device__ __forceinline__ void run(const WafHorizontalFused &a, int bx, int by, int bz) {
    waf_h_body<float, double, S, TR, PER>(a.f, &a.ap2, &a.wzp2, a.nf2, bx, by, bz);
  }
};

// The mixed form's two field batches in one launch: the float fields and
// the double fields (pai, tetav) of a WAF stage.  Blocks whose index falls
// in the first range run the float batch, the others the double batch, so
// the small double batch no longer needs a launch (and its tail) of its own.
template <typename S, int GF, int GD>
struct WafVerticalMixed {
  WafVertical<float, S, GF> f;
  WafVertical<double, S, GD> p;
  int ngf;                     // blocks with bz < ngf run the float batch
  static __device__ __forceinline__ void run(const WafVerticalMixed &a, int bx, int by, int bz) {
    if (bz < a.ngf)
      WafVertical<float, S, GF>::run(a.f, bx, by, bz);
    else
      WafVertical<double, S, GD>::run(a.p, bx, by, bz - a.ngf);
  }
};

template <typename S, int TR, int PER>
struct WafHorizontalMixed {
  WafHorizontal<float, S, TR, PER> f;
  WafHorizontal<double, S, TR, PER> p;
  int gxf;                     // blocks with bx < gxf run the float batch
  static __device__ __forceinline__ void run(const WafHorizontalMixed &a, int bx, int by, int bz) {
    if (bx < a.gxf)
      WafHorizontal<float, S, TR, PER>::run(a.f, bx, by, bz);
    else
      WafHorizontal<double, S, TR, PER>::run(a.p, bx - a.gxf, by, bz);
  }
};

// ---------------------------------------------------------------------------
// The same meridional + zonal passes in row-streaming form.  A wavefront
// owns 64 consecutive j points of one level (kWrOut outputs and 2 halo
// points each side) and marches down a chunk of kWrRows i rows.  Each lane
// keeps the last four rows of wz in registers, so the meridional flux needs
// no other lane; the zonal neighbours of p0 and of the zonal flux come from
// the other lanes of the wavefront by shuffles.  No LDS, no barriers.  A
// lane handles a group of up to kWrG fields, so the field-independent
// coefficients of a point are formed once per group.  Every value is formed
// with the operations and operand order of WafHorizontal.
constexpr int kWrOut = 60;     // output j points per wavefront
constexpr int kWrWaves = 4;    // wavefronts (i chunks) per block

template <typename T>
__device__ __forceinline__ T sel4(int o, T a0, T a1, T a2, T a3) {
  return o == 0 ? a0 : (o == 1 ? a1 : (o == 2 ? a2 : a3));
}

template <typename T, typename S, int kWrG>
struct WafRows {
  Dims d;
  AdvPtrs<T> ap;
  AdvPtrs<T> wzp;
  int nf;
  int ngroups;
  int rows;                    // i rows per chunk
  const S *u, *v, *mu, *mv, *mx, *mx2, *rmu, *rmv, *rfmzu, *rfmzv, *fmz;
  double dtrdx_, dtrdy_;
  int rotllr;
  static __device__ __forceinline__ void run(const WafRows &args_, int bx, int by, int bz);
};

template <typename T, typename S, int kWrG>
__device__ __forceinline__ void WafRows<T, S, kWrG>::run(const WafRows &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const int jce2 = d.jx - 1, ice2 = d.iy - 1;
  const int jmin = 1, jmax = jce2, imin = 1, imax = ice2;
  const int rows = args_.rows;
  const int chunk = by * kWrWaves + (int)threadIdx.y;
  const int i0 = 1 + chunk * rows;
  if (i0 > ice2) return;                          // uniform per wavefront
  const int i1 = i0 + rows - 1 < ice2 ? i0 + rows - 1 : ice2;
  const int ngroups = args_.ngroups;
  const int grp = bz % ngroups;
  const int k = 1 + bz / ngroups;
  const int f0 = grp * kWrG;
  const int ng = args_.nf - f0 < kWrG ? args_.nf - f0 : kWrG;
  const int lane = threadIdx.x;
  const int jj = bx * kWrOut - 1 + lane;          // j0 - 2 + lane, j0 = 1 + kWrOut*bx
  const bool jin = jj >= 1 && jj <= jce2;
  const bool jout = jin && lane >= 2 && lane < 2 + kWrOut;
  const int jc = jin ? jj : 1;
  const int rotllr = args_.rotllr;
  const S dtrdx = S(args_.dtrdx_), dtrdy = S(args_.dtrdy_);
  const long long pj = d.pj;
  const S *__restrict__ u = args_.u;
  const S *__restrict__ v = args_.v;
  const S *__restrict__ mu = args_.mu;
  const S *__restrict__ mv = args_.mv;
  const S *__restrict__ mx = args_.mx;
  const S *__restrict__ mx2 = args_.mx2;
  const S *__restrict__ rmu = args_.rmu;
  const S *__restrict__ rmv = args_.rmv;
  const S *__restrict__ rfmzu = args_.rfmzu;
  const S *__restrict__ rfmzv = args_.rfmzv;
  const S *__restrict__ fmz = args_.fmz;
  const T *wz[kWrG];
  const T *pp[kWrG];
  T *po[kWrG];
#pragma unroll
  for (int g = 0; g < kWrG; ++g) {
    const int f = f0 + (g < ng ? g : 0);
    wz[g] = args_.wzp.src[f];
    pp[g] = args_.ap.src[f];
    po[g] = args_.ap.dst[f];
  }
  auto wz_row = [&](int g, int r) -> T {
    return (jin && r >= 1 && r <= ice2) ? wz[g][d.at(jj, r, k)] : T(0);
  };
  // meridional flux through face ii from the window rows ii-2 .. ii+1
  auto zpby = [&](int ii, S vv, S mvv, const T *W) -> T {
    const S zamu = rotllr ? vv * dtrdy : vv * mvv * dtrdy;
    T is;
    int ih;
    if (zamu > S(0)) {
      is = T(1);
      ih = ii - 1;
    } else {
      is = T(-1);
      ih = ii + 1 < imax ? ii + 1 : imax;
    }
    const int ihm1 = ih - 1 > imin ? ih - 1 : imin;
    const T wh = sel4(ih - (ii - 2), W[0], W[1], W[2], W[3]);
    const T whm1 = sel4(ihm1 - (ii - 2), W[0], W[1], W[2], W[3]);
    const T rr = flow_param(wh - whm1, W[2] - W[1]);
    const T b = waf_b(rr);
    const T zphi = is + T(zamu) * b - is * b;
    return T(S(0.5) * vv) * ((T(1) + zphi) * W[1] + (T(1) - zphi) * W[2]);
  };
  auto face_v = [&](int ii, S *vv, S *mvv) {
    const bool ok = jin && ii >= 2 && ii <= ice2;
    *vv = ok ? v[d.at(jc, ii, k)] : S(0);
    *mvv = ok ? mv[d.at2(jc, ii)] : S(0);
  };

  T W[kWrG][4], zprev[kWrG];
  S v_i, mv_i;                                    // v, mv on face i
  face_v(i0, &v_i, &mv_i);
#pragma unroll
  for (int g = 0; g < kWrG; ++g) {
    if (g < ng) {
      T W0[4] = {wz_row(g, i0 - 2), wz_row(g, i0 - 1), wz_row(g, i0), wz_row(g, i0 + 1)};
      zprev[g] = (jin && i0 >= 2) ? zpby(i0, v_i, mv_i, W0) : T(0);
      W[g][0] = W0[1];
      W[g][1] = W0[2];
      W[g][2] = W0[3];
      W[g][3] = wz_row(g, i0 + 2);
    }
  }
  for (int i = i0; i <= i1; ++i) {
    S v_n, mv_n;                                  // face i+1
    face_v(i + 1, &v_n, &mv_n);
    const bool row_in = i >= 2 && i <= ice2 - 1;
    const long long c = d.at(jc, i, k), c2 = d.at2(jc, i);
    // field-independent coefficients of p0 (row i) and of the zonal pass
    S zrfmn = S(0), zrfms = S(0), cdv = S(0), mf = S(1);
    S uu = S(0), zamu_u = S(0), is_u = S(0);
    int oh = 0, ohm1 = 0;
    if (jin && row_in) {
      if (rotllr) {
        const S zhxvtn = dtrdy * rmv[c2 + pj] * mx[c2];
        const S zhxvts = dtrdy * rmv[c2] * mx[c2];
        zrfmn = zhxvtn * fmz[c] * rfmzv[c + pj];
        zrfms = zhxvts * fmz[c] * rfmzv[c];
        cdv = v[c + pj] * zrfmn - v[c] * zrfms;
      } else {
        // rfmzu, not rfmzv, on the meridional fluxes (:1004-1005)
        zrfmn = dtrdy * fmz[c] * rfmzu[c + pj];
        zrfms = dtrdy * fmz[c] * rfmzu[c];
        cdv = v[c + pj] * rmv[c2 + pj] * zrfmn - v[c] * rmv[c2] * zrfms;
        mf = mx2[c2];
      }
      if (jj >= 2) {
        uu = u[c];
        zamu_u = uu * mu[c2] * dtrdx;
        int jh;
        if (zamu_u > S(0)) {
          is_u = S(1);
          jh = jj - 1;
        } else {
          is_u = S(-1);
          jh = jj + 1 < jmax ? jj + 1 : jmax;
        }
        const int jhm1 = jh - 1 > jmin ? jh - 1 : jmin;
        oh = jh - jj + 2;                         // offsets -2..+1 -> 0..3
        ohm1 = jhm1 - jj + 2;
      }
    }
    // coefficients of the final update (interior points of row i)
    const bool o_int = jout && row_in && jj >= 2 && jj <= jce2 - 1;
    S o_zrfmw = S(0), o_zrfme = S(0), o_cdv = S(0), o_mf = S(1);
    if (o_int) {
      if (rotllr) {
        const S zcostx = dtrdx * mx[c2];
        o_zrfmw = zcostx * fmz[c] * rfmzu[c];
        o_zrfme = zcostx * fmz[c] * rfmzu[c + 1];
        o_cdv = u[c + 1] * o_zrfme - u[c] * o_zrfmw;
      } else {
        o_zrfmw = dtrdx * fmz[c] * rfmzu[c];
        o_zrfme = dtrdx * fmz[c] * rfmzu[c + 1];
        o_cdv = u[c + 1] * rmu[c2 + 1] * o_zrfme - u[c] * rmu[c2] * o_zrfmw;
        o_mf = mx2[c2];
      }
    }
#pragma unroll
    for (int g = 0; g < kWrG; ++g) {
      if (g < ng) {
        const T znext = (jin && i + 1 <= ice2) ? zpby(i + 1, v_n, mv_n, W[g]) : T(0);
        const T pold = jin ? pp[g][c] : T(0);
        T p0 = T(0);
        if (jin && row_in) {
          const T zdv = T(cdv) * pold;
          const T x = zprev[g] * T(zrfms) - znext * T(zrfmn) + zdv;
          // ROTLLR has no mx2 factor; mf = 1 there and 1*x == x exactly
          p0 = W[g][1] + T(mf) * x;
        }
        // zonal flux through face jj from p0 at jj-2 .. jj+1
        const T pm2 = __shfl(p0, lane >= 2 ? lane - 2 : 0);
        const T pm1 = __shfl(p0, lane >= 1 ? lane - 1 : 0);
        const T pp1 = __shfl(p0, lane <= 62 ? lane + 1 : 63);
        T zpbw = T(0);
        if (jin && row_in && jj >= 2) {
          const T rr = flow_param(sel4(oh, pm2, pm1, p0, pp1) - sel4(ohm1, pm2, pm1, p0, pp1),
                                  p0 - pm1);
          const T b = waf_b(rr);
          const T zphi = T(is_u) + T(zamu_u) * b - T(is_u) * b;
          zpbw = T(S(0.5) * uu) * ((T(1) + zphi) * pm1 + (T(1) - zphi) * p0);
        }
        const T ze = __shfl(zpbw, lane <= 62 ? lane + 1 : 63);
        if (jout) {
          T val = pold;
          if (o_int) {
            const T zdv = T(o_cdv) * pold;
            if (rotllr)
              val = p0 + zpbw * T(o_zrfmw) - ze * T(o_zrfme) + zdv;
            else
              val = p0 + T(o_mf) * (zpbw * T(o_zrfmw) - ze * T(o_zrfme) + zdv);
          }
          po[g][c] = val;
        }
        zprev[g] = znext;
        W[g][0] = W[g][1];
        W[g][1] = W[g][2];
        W[g][2] = W[g][3];
        W[g][3] = wz_row(g, i + 3);
      }
    }
    v_i = v_n;
    mv_i = mv_n;
  }
}

// ---------------------------------------------------------------------------
// Curvature terms (:811-825), pointwise on interior cross points.
template <typename R>
struct Curvature {
  Dims d;
  R * ux;
  R * vx;
  const R * mu;
  const R * mv;
  const R * tanrot;
  double rdx_;
  double dta_;
  int rotllr;
  static __device__ __forceinline__ void run(const Curvature &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void Curvature<R>::run(const Curvature &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  R *__restrict__ ux = args_.ux;
  R *__restrict__ vx = args_.vx;
  const R *__restrict__ mu = args_.mu;
  const R *__restrict__ mv = args_.mv;
  const R *__restrict__ tanrot = args_.tanrot;
  const double rdx_ = args_.rdx_;
  const double dta_ = args_.dta_;
  const int rotllr = args_.rotllr;

  const int j = d.jc1 + bx * TJ + threadIdx.x;
  const int i = d.ic1 + by * TI + threadIdx.y;
  const int k = 1 + bz;
  if (j > d.jc2 || i > d.ic2) return;
  const R rdx = R(rdx_), dta = R(dta_);
  const long long c = d.at(j, i, k), c2 = d.at2(j, i);
  R tanx, tany;
  if (rotllr) {
    tanx = tanrot[c2];
    tany = tanx;
  } else {
    tanx = (mu[c2 + d.oj(j, -1)] - mu[c2]) * rdx;
    tany = (mv[c2 + d.oi(i, -1)] - mv[c2]) * rdx;
  }
  R a = ux[c];
  const R b = vx[c];
  a = a + a * b * tanx * dta;
  ux[c] = a;
  vx[c] = b - a * a * tany * dta;
}

// ---------------------------------------------------------------------------
// ux,vx -> u,v (:1477-1522) and, when do_w, wx -> w (:1461-1475).
template <typename R>
struct StaggerOut {
  Dims d;
  const R * ux;
  const R * vx;
  const R * wx;
  R * u;
  R * v;
  double * w;
  int do_w;
  int kc;            // levels per thread: levels 1 + bz*kc .. of the column
  static __device__ __forceinline__ void run(const StaggerOut &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void StaggerOut<R>::run(const StaggerOut &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const R *__restrict__ ux = args_.ux;
  const R *__restrict__ vx = args_.vx;
  const R *__restrict__ wx = args_.wx;
  R *__restrict__ u = args_.u;
  R *__restrict__ v = args_.v;
  double *__restrict__ w = args_.w;
  const int do_w = args_.do_w;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  const int k0 = 1 + bz * args_.kc;
  const int k1 = k0 + args_.kc - 1 < kz ? k0 + args_.kc - 1 : kz;
  if (j > d.je2 || i > d.ie2) return;
  const long long pj = d.pj, P = d.plane;
  const R h = R(0.5), a9 = R(0.5625), a1 = R(0.0625);
  // Every load of a level first, at neighbour offsets clamped into the
  // domain where a branch below does not use them (j + 1 <= jx and
  // i + 1 <= iy hold here; a periodic domain wraps), all waited for once:
  // otherwise a later field's loads follow an earlier field's store, and
  // one wait counter covers both.
  const int crm = d.crm;
  const long long dw1 = crm ? d.oj(j, -1) : -(j >= 2 ? 1 : 0);
  const long long dw2 = crm ? d.oj(j, -2) : -(j >= 3 ? 2 : 0);
  const long long de1 = crm ? d.oj(j, 1) : 1;
  const long long ds1 = crm ? d.oi(i, -1) : -(i >= 2 ? pj : 0);
  const long long ds2 = crm ? d.oi(i, -2) : -(i >= 3 ? 2 * pj : 0);
  const long long dn1 = crm ? d.oi(i, 1) : pj;
  const bool du = !crm && j >= 2 && i >= 2 && i <= iy - 2;
  const bool dv = !crm && i >= 2 && j >= 2 && j <= jx - 2;
  for (int k = k0; k <= k1; ++k) {
    const long long c = d.at(j, i, k);
    const R u0 = ux[c], uw = ux[c + dw1], ue = ux[c + de1], uww = ux[c + dw2];
    const R v0 = vx[c], vs = vx[c + ds1], vn = vx[c + dn1], vss = vx[c + ds2];
    R w0 = R(0), wb = R(0), wa = R(0), wbb = R(0);
    if (do_w) {
      w0 = wx[c];
      wb = wx[c - (k >= 2 ? P : 0)];
      wa = wx[c + (k < kz ? P : 0)];
      wbb = wx[c - (k >= 3 ? 2 * P : 0)];
    }
    issued(u0, uw, ue, uww);
    issued(v0, vs, vn, vss);
    issued(w0, wb, wa, wbb);
    if (crm) {
      // periodic: the fourth-order interpolation everywhere, neighbours
      // wrapped (no has_bdy* edge)
      u[c] = a9 * (u0 + uw) - a1 * (ue + uww);
      v[c] = a9 * (v0 + vs) - a1 * (vn + vss);
    }
    if (du) {
      R a;
      if (j == 2)
        a = h * (u0 + uw);
      else if (j == jx - 1)
        a = h * (uw + u0);
      else
        a = a9 * (u0 + uw) - a1 * (ue + uww);
      u[c] = a;
    }
    if (dv) {
      R a;
      if (i == 2)
        a = h * (v0 + vs);
      else if (i == iy - 1)
        a = h * (vs + v0);
      else
        a = a9 * (v0 + vs) - a1 * (vn + vss);
      v[c] = a;
    }
    if (do_w && k >= 2) {
      R a;
      if (k == 2)
        a = h * (w0 + wb);
      else if (k == kz)
        a = h * (w0 + wb);
      else
        a = a9 * (w0 + wb) - a1 * (wa + wbb);
      w[c] = (double)a;
    }
  }
}

// ---------------------------------------------------------------------------
// Lateral boundaries of a real-case run: boundary() (:448-529).
//
// Boundary data (b0/b1 of u, v, t, pai, qv, qc, qi and ps at the two
// bracketing times) and the Davies weights are double in both forms.
struct LbcData {
  const double *ps0, *ps1, *u0, *u1, *v0, *v1, *t0, *t1, *pai0, *pai1;
  const double *qv0, *qv1, *qc0, *qc1, *qi0, *qi1;
  const double *xfu, *xfv, *xfc, *tnudge;
};

// qxzeroval, qxcheckval (Main/mpplib/mod_runparams.F90:184-192)
__device__ __forceinline__ double qx_zero(int n) {
  return n == 0 ? 1.0e-8 : (n < 7 ? 0.0 : (n == 7 ? 1.0e10 : (n == 8 ? 100.0 : 0.01)));
}
__device__ __forceinline__ double qx_check(int n) {
  return n == 0 ? 1.0e-8 : (n < 7 ? 1.0e-16 : (n == 7 ? 1.0e10 : (n == 8 ? 100.0 : 0.01)));
}

// bdyval, MOLOCH branch (Main/mod_bdycod.F90:1638-1866).  sn = 0: the west
// and east columns (corners excluded), one thread per (i, k); sn = 1: the
// south and north rows (corners included), one thread per (j, k).  The two
// run as separate launches in that order, as the reference's sequence needs
// (a corner's outflow copy reads the value the west/east pass wrote).
template <typename R>
struct BdyVal {
  Dims d;
  LbcData b;
  R *u, *v, *t;
  double *pai, *w;
  QPtrs<R> q;
  R *ps;
  double x1;
  int nqx, ipptls, present_qc, present_qi, sn;
  R *tke;            // ibltyp 2, else null
  double tkemin;
  static __device__ __forceinline__ void run(const BdyVal &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void BdyVal<R>::run(const BdyVal &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const LbcData &b = args_.b;
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  const int n_along = args_.sn ? jx : iy;
  const int m = bx * (int)blockDim.x + (int)threadIdx.x;     // j or i
  const int k = 1 + by;
  if (m < 1 || m > n_along || k > kz) return;
  const double x1 = args_.x1;
  const double x0 = 1.0 - x1;
  const int pqc = args_.present_qc, pqi = args_.present_qi;
  const int nqx = args_.nqx;
  R *__restrict__ u = args_.u;
  R *__restrict__ v = args_.v;
  R *__restrict__ t = args_.t;
  double *__restrict__ pai = args_.pai;
  double *__restrict__ w = args_.w;
  R *__restrict__ ps = args_.ps;
  auto lin = [&](const double *b0, const double *b1, long long c) {
    return x0 * b0[c] + x1 * b1[c];
  };
  for (int side = 0; side < 2; ++side) {
    if (!args_.sn) {
      const int i = m;
      const int jd = side == 0 ? 1 : jx;
      const int jc = side == 0 ? 1 : jx - 1;
      const int jn = side == 0 ? 2 : jx - 2;
      const bool ici = i >= 2 && i <= iy - 2, idi = i >= 2 && i <= iy - 1;
      const long long cd = d.at(jd, i, k), cc = d.at(jc, i, k), cn = d.at(jn, i, k);
      if (ici && k == 1) ps[d.at2(jc, i)] = R(x0 * b.ps0[d.at2(jc, i)] + x1 * b.ps1[d.at2(jc, i)]);
      if (ici) u[cd] = R(lin(b.u0, b.u1, cd));
      if (idi) v[cc] = R(lin(b.v0, b.v1, cc));
      if (ici) {
        t[cc] = R(lin(b.t0, b.t1, cc));
        pai[cc] = lin(b.pai0, b.pai1, cc);
        args_.q.q[0][cc] = R(lin(b.qv0, b.qv1, cc));
        if (pqc) args_.q.q[1][cc] = R(lin(b.qc0, b.qc1, cc));
        if (pqi && args_.ipptls > 1) args_.q.q[2][cc] = R(lin(b.qi0, b.qi1, cc));
        const double uu = u[cd];
        const bool inflow = side == 0 ? uu > 0.0 : uu < 0.0;
        // over the fixed maximum (10 species) with a guard: constant
        // indices keep the argument struct out of scratch memory
#pragma unroll
        for (int n = 2; n <= 10; ++n) {
          if (n > nqx || (pqc && n == 2) || (pqi && n == 3)) continue;
          const R qxint = args_.q.q[n - 1][cn];
          args_.q.q[n - 1][cc] = inflow ? R(qx_zero(n - 1)) : qxint;
        }
        w[cc] = inflow ? 0.0 : w[cn];
        if (args_.tke) {
          // TKE rows (Main/mod_bdycod.F90:1684-1694, :1742-1752): level 1 at
          // tkemin; level k+1 (k = 2..kz) from the inflow sign of u(k) + u(k-1),
          // u(k-1) being the value this stage writes at that level
          R *tke = args_.tke;
          if (k == 1) {
            tke[cc] = R(args_.tkemin);
          } else {
            const double us = (double)u[cd] + (double)R(lin(b.u0, b.u1, cd - d.plane));
            const bool tin = side == 0 ? us > 0.0 : us < 0.0;
            tke[cc + d.plane] = tin ? R(args_.tkemin) : tke[cn + d.plane];
          }
        }
      }
    } else {
      const int j = m;
      const int id = side == 0 ? 1 : iy;
      const int ic = side == 0 ? 1 : iy - 1;
      const int in = side == 0 ? 2 : iy - 2;
      const bool jce = j <= jx - 1;
      const long long cdu = d.at(j, ic, k);          // u row: j in jde at i = ice
      const long long cdv = d.at(j, id, k), cc = d.at(j, ic, k), cn = d.at(j, in, k);
      if (jce && k == 1) ps[d.at2(j, ic)] = R(x0 * b.ps0[d.at2(j, ic)] + x1 * b.ps1[d.at2(j, ic)]);
      u[cdu] = R(lin(b.u0, b.u1, cdu));
      if (jce) {
        v[cdv] = R(lin(b.v0, b.v1, cdv));
        t[cc] = R(lin(b.t0, b.t1, cc));
        pai[cc] = lin(b.pai0, b.pai1, cc);
        args_.q.q[0][cc] = R(lin(b.qv0, b.qv1, cc));
        if (pqc) args_.q.q[1][cc] = R(lin(b.qc0, b.qc1, cc));
        if (pqi && args_.ipptls > 1) args_.q.q[2][cc] = R(lin(b.qi0, b.qi1, cc));
        const double vv = v[cdv];
        const bool inflow = side == 0 ? vv > 0.0 : vv < 0.0;
        // over the fixed maximum (10 species) with a guard: constant
        // indices keep the argument struct out of scratch memory
#pragma unroll
        for (int n = 2; n <= 10; ++n) {
          if (n > nqx || (pqc && n == 2) || (pqi && n == 3)) continue;
          const R qxint = args_.q.q[n - 1][cn];
          args_.q.q[n - 1][cc] = inflow ? R(qx_zero(n - 1)) : qxint;
        }
        w[cc] = inflow ? 0.0 : w[cn];
        if (args_.tke) {
          // (Main/mod_bdycod.F90:1800-1810, :1858-1868), j over jce
          R *tke = args_.tke;
          if (k == 1) {
            tke[cc] = R(args_.tkemin);
          } else {
            const double vs = (double)v[cdv] + (double)R(lin(b.v0, b.v1, cdv - d.plane));
            const bool tin = side == 0 ? vs > 0.0 : vs < 0.0;
            tke[cc + d.plane] = tin ? R(args_.tkemin) : tke[cn + d.plane];
          }
        }
      }
    }
  }
}

// motopnudge (Main/mod_bdycod.F90:4049-4081, wfac = 0, dta = dtsec) then
// the Davies relaxation of boundary() (:476-495): morelax_external on u, v,
// t, pai, qv, qc, qi and morelax_fraction(frac = 0) on w
// (Main/mod_bdycod.F90:3962-4031), one thread per point, in the reference's
// order; every statement of those loops touches only its own point.
template <typename R>
struct BdyRelax {
  Dims d;
  LbcData b;
  R *u, *v, *t;
  double *pai, *w;
  QPtrs<R> q;
  double x1, dta;
  int ipptls, present_qc, present_qi, top_nudge, nztop;
  static __device__ __forceinline__ void run(const BdyRelax &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void BdyRelax<R>::run(const BdyRelax &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const LbcData &b = args_.b;
  const int j = 2 + bx * TJ + threadIdx.x;
  const int i = 2 + by * TI + threadIdx.y;
  const int k = 1 + bz;
  const int jx = d.jx, iy = d.iy;
  if (j > jx - 1 || i > iy - 1) return;
  const long long c = d.at(j, i, k);
  const double x1 = args_.x1;
  const double x0 = 1.0 - x1;
  const bool jci = j <= jx - 2, ici = i <= iy - 2;   // j >= 2, i >= 2 by the grid
  const bool u_pt = ici, v_pt = jci, c_pt = jci && ici;
  R *__restrict__ u = args_.u;
  R *__restrict__ v = args_.v;
  R *__restrict__ t = args_.t;
  double *__restrict__ pai = args_.pai;
  double *__restrict__ w = args_.w;
  if (args_.top_nudge) {
    const double trtau = 1.0 / (2.0 * 3600.0);
    const double wrtau = 1.0 / (3600.0 / 4.0);
    const double dta = args_.dta;
    if (c_pt && k == 2) {
      const double fext = 0.0 * w[c];
      const double xf = wrtau * dta;
      w[c] = (1.0 - xf) * w[c] + xf * fext;
    }
    if (k <= args_.nztop) {
      const double xf = b.tnudge[k - 1] * trtau * dta;
      if (c_pt) t[c] = R((1.0 - xf) * (double)t[c] + xf * (x0 * b.t0[c] + x1 * b.t1[c]));
      if (u_pt) u[c] = R((1.0 - xf) * (double)u[c] + xf * (x0 * b.u0[c] + x1 * b.u1[c]));
      if (v_pt) v[c] = R((1.0 - xf) * (double)v[c] + xf * (x0 * b.v0[c] + x1 * b.v1[c]));
    }
  }
  auto relax = [&](double fv, double xf, const double *b0, const double *b1) {
    const double fext = (x0 * b0[c] + x1 * b1[c]);
    return (1.0 - xf) * fv + xf * fext;
  };
  if (u_pt && b.xfu[c] >= 0.0) u[c] = R(relax(u[c], b.xfu[c], b.u0, b.u1));
  if (v_pt && b.xfv[c] >= 0.0) v[c] = R(relax(v[c], b.xfv[c], b.v0, b.v1));
  if (c_pt) {
    const double xf = b.xfc[c];
    if (xf >= 0.0) {
      t[c] = R(relax(t[c], xf, b.t0, b.t1));
      pai[c] = relax(pai[c], xf, b.pai0, b.pai1);
      args_.q.q[0][c] = R(relax(args_.q.q[0][c], xf, b.qv0, b.qv1));
      w[c] = (1.0 - xf) * w[c] + xf * w[c] * 0.0;
      if (args_.ipptls > 0) {
        if (args_.present_qc) args_.q.q[1][c] = R(relax(args_.q.q[1][c], xf, b.qc0, b.qc1));
        if (args_.ipptls > 1 && args_.present_qi)
          args_.q.q[2][c] = R(relax(args_.q.q[2][c], xf, b.qi0, b.qi1));
      }
    }
  }
}

// Spectral nudging of one field (mospectral_nudge, Main/mod_bdycod.F90:
// 3940-3960, with lowpass_filter :3898-3938 on a single domain).  Stage 0:
// zn = (x0 b0 + x1 b1) - f over j1..j2 x i1..i2; 1: sx(i,m) = sum over
// j = jj1..jj2 of zn(j,i) bvx(j,m); 2: zn = sum over m of sx(i,m) bvx(j,m);
// 3: sy(j,l) = sum over i = ii1..ii2 of zn(j,i) bvy(i,l); 4: f += cnudge *
// sum over l of sy(j,l) bvy(i,l) on jj1..jj2 x ii1..ii2.  Every sum runs in
// the reference's order, one thread per output; all levels in one launch.
template <typename R>
struct SpecNudge {
  Dims d;
  R *f;
  const double *b0, *b1, *bvx, *bvy, *cnudge;
  double *zn, *sx, *sy;
  double x1;
  int j1, j2, i1, i2, jj1, jj2, ii1, ii2, km2, lm2, stage;
  static __device__ __forceinline__ void run(const SpecNudge &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SpecNudge<R>::run(const SpecNudge &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const SpecNudge &a = args_;
  const int k = 1 + bz;
  const int jx = d.jx, iy = d.iy;
  const int x = bx * (int)blockDim.x + (int)threadIdx.x;      // j, or the i / j of a sum
  const int y = by;                                            // i, or the mode
  const long long zk = (long long)(k - 1) * d.plane;
  switch (a.stage) {
    case 0: {
      const int j = x + 1, i = y + 1;
      if (j < a.j1 || j > a.j2 || i < a.i1 || i > a.i2) return;
      const long long c = d.at(j, i, k);
      const double x0 = 1.0 - a.x1;
      a.zn[c] = (x0 * a.b0[c] + a.x1 * a.b1[c]) - (double)a.f[c];
      return;
    }
    case 1: {
      const int i = x + 1, m = y + 1;
      if (i < a.i1 || i > a.i2 || m > a.km2) return;
      double acc = 0.0;
      for (int j = a.jj1; j <= a.jj2; ++j)
        acc = acc + a.zn[zk + d.at2(j, i)] * a.bvx[(long long)(m - 1) * jx + (j - 1)];
      a.sx[((long long)(k - 1) * a.km2 + (m - 1)) * iy + (i - 1)] = acc;
      return;
    }
    case 2: {
      const int j = x + 1, i = y + 1;
      if (j < a.j1 || j > a.j2 || i < a.i1 || i > a.i2) return;
      double acc = 0.0;
      for (int m = 1; m <= a.km2; ++m)
        acc = acc + a.sx[((long long)(k - 1) * a.km2 + (m - 1)) * iy + (i - 1)] *
                        a.bvx[(long long)(m - 1) * jx + (j - 1)];
      a.zn[zk + d.at2(j, i)] = acc;
      return;
    }
    case 3: {
      const int j = x + 1, l = y + 1;
      if (j < a.j1 || j > a.j2 || l > a.lm2) return;
      double acc = 0.0;
      for (int i = a.ii1; i <= a.ii2; ++i)
        acc = acc + a.zn[zk + d.at2(j, i)] * a.bvy[(long long)(l - 1) * iy + (i - 1)];
      a.sy[((long long)(k - 1) * a.lm2 + (l - 1)) * jx + (j - 1)] = acc;
      return;
    }
    default: {
      const int j = x + 1, i = y + 1;
      if (j < a.jj1 || j > a.jj2 || i < a.ii1 || i > a.ii2) return;
      double acc = 0.0;
      for (int l = 1; l <= a.lm2; ++l)
        acc = acc + a.sy[((long long)(k - 1) * a.lm2 + (l - 1)) * jx + (j - 1)] *
                        a.bvy[(long long)(l - 1) * iy + (i - 1)];
      const long long c = d.at(j, i, k);
      a.f[c] = R((double)a.f[c] + a.cnudge[k - 1] * acc);
      return;
    }
  }
}

// mkslice, MOLOCH branch (Main/mod_slice.F90:112-160, 340-382).  A block
// holds 64 columns of one row i (lane = j) and kSliceWaves wavefronts that
// split the levels, so every level's pointwise work (pf, th3d, the qx floor,
// rhb3d, wpx3d) is one coalesced pass.  The three column searches are
// reductions over the levels: ktrop is the lowest level of 2..kz-1 below
// ptrop (the reference walks up from kz-1 and stops there), kmxpbl is one
// above the lowest level of 2..kz-1 higher than 5 km, and the 700 hPa
// bracket is the highest level of 2..kz-1 below 700 hPa; each wave keeps its
// extremum and wave 0 combines them through LDS and writes the 2-D slices.
// th3d at the bracket is recomputed with the same expression, so th700 has
// the same bits as the reference's read-back.
constexpr int kSliceWaves = 4;

template <typename R>
struct Slice {
  Dims d;
  const double *pai, *w, *zetaf, *xlat;
  const R *t, *p, *rho, *qsat, *zeta, *ps;
  QPtrs<R> q;
  double *pf, *th, *rh, *omega, *rhox, *tp, *th700, *ptrop, *ktrop, *kmxpbl;
  double rhmin, rhmax, calday, dayspy;
  int nqx, icldmstrat, irceideal, ibltyp;
  static __device__ __forceinline__ void run(const Slice &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void Slice<R>::run(const Slice &args_, int bx, int by, int bz) {
  const Slice &a = args_;
  const Dims d = a.d;
  const int lane = threadIdx.x, wv = threadIdx.y;
  const int j = 1 + bx * 64 + lane;
  const int i = 1 + by;
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  const bool col = j <= jx && i <= iy;
  const bool ce = j <= d.je2 && i <= d.ie2;                          // jce x ice
  const bool inner = j >= d.jc1 && j <= d.jc2 && i >= d.ic1 && i <= d.ic2;  // jci x ici
  const long long c2 = col ? d.at2(j, i) : 0, P = d.plane;
  const double rovcp = kRgas * (1.0 / kCpd);
  __shared__ double s_ptrop[64];
  __shared__ int s_red[3][kSliceWaves][64];
  // tropopause pressure (Mateus, Mendes and Pires 2022), once per column
  if (wv == 0) {
    double pt = 0.0;
    if (inner) {
      if (a.irceideal != 1) {
        // anorth / asouth fits (Main/mod_slice.F90:343-360)
        const double lat = a.xlat[c2];
        const bool nh = lat > 0.0;
        const double c0 = nh ? 7.9925 : 8.1797, c1 = nh ? 8.3329 : 8.1455,
                     c2f = nh ? 24.1731 : -23.4839, c3 = nh ? -1.8069 : 1.1464,
                     c4 = nh ? 0.1082 : 0.0798, c5 = nh ? -0.1493 : -0.1491;
        const double twopi = kMathpi * 2.0;
        const double ztrop = c0 + c1 / pow(1.0 + exp(-(lat - c2f) / c3), c4) +
                             c5 * cos((twopi * (a.calday - 28.0)) / a.dayspy);
        pt = kP00 * exp(-ztrop / 8.4);
        a.ptrop[c2] = pt;
      } else {
        pt = a.ptrop[c2];
      }
    }
    s_ptrop[lane] = pt;
  }
  __syncthreads();
  int kt = 0, kp = 0, k7 = kz + 1;
  if (ce) {
    const double ptr = s_ptrop[lane];
    for (int k = 1 + wv; k <= kz; k += kSliceWaves) {
      const long long c = d.at(j, i, k);
      if (k >= 2) {
        a.pf[c] = kP00 * pow(0.5 * (a.pai[c] + a.pai[c - P]), kCpovr);
      } else {
        const double pf2 = kP00 * pow(0.5 * (a.pai[c + P] + a.pai[c]), kCpovr);
        a.pf[c] = pf2 - kEgrav * (double)a.rho[c] * (a.zetaf[c] - a.zetaf[c + P]);
        a.pf[c + (long long)kz * P] = (double)a.ps[c2];
      }
      const double pk = (double)a.p[c];
      a.th[c] = (double)a.t[c] * pow(kP00 / pk, rovcp);
      if (!inner) continue;
#pragma unroll
      for (int n = 0; n < 10; ++n)
        if (n < a.nqx && a.q.q[n][c] < R(qx_check(n))) a.q.q[n][c] = R(qx_zero(n));
      const double r = (double)a.q.q[0][c] / (double)a.qsat[c];
      a.rh[c] = fmin(fmax(r, a.rhmin), a.rhmax);
      a.omega[c] = -kEgrav * (double)a.rho[c] * 0.5 * (a.w[c + P] + a.w[c]);
      if (k >= 2 && k <= kz - 1) {
        if (pk < ptr && k > kt) kt = k;
        if (a.ibltyp == 1 && (double)a.zeta[c] > 5000.0 && k > kp) kp = k;
        if (a.icldmstrat == 1 && pk > 70000.0 && k < k7) k7 = k;
      }
    }
  }
  s_red[0][wv][lane] = kt;
  s_red[1][wv][lane] = kp;
  s_red[2][wv][lane] = k7;
  __syncthreads();
  if (wv != 0 || !col) return;
  if (!inner) {                                                      // frame
    a.ktrop[c2] = kz;
    if (a.ibltyp == 1) a.kmxpbl[c2] = kz;
    return;
  }
  for (int w2 = 1; w2 < kSliceWaves; ++w2) {
    kt = max(kt, s_red[0][w2][lane]);
    kp = max(kp, s_red[1][w2][lane]);
    k7 = min(k7, s_red[2][w2][lane]);
  }
  const bool search = kz - 1 >= 2;
  a.ktrop[c2] = !search ? kz : (kt ? kt : 2);
  if (a.ibltyp == 1) a.kmxpbl[c2] = !search ? kz : (kp ? kp + 1 : 2);
  const long long cz = d.at(j, i, kz);
  const double psv = (double)a.ps[c2];
  a.rhox[c2] = psv / (kRgas * (double)a.t[cz]);
  a.tp[c2] = (double)a.t[cz] * pow(psv / (double)a.p[cz], rovcp);
  if (a.icldmstrat == 1) {
    double v;
    if (k7 <= kz - 1) {
      const long long c1 = d.at(j, i, k7), c0 = c1 - P;
      const double p1 = (double)a.p[c1], p0 = (double)a.p[c0];
      const double w1 = (p1 - 70000.0) / (p1 - p0);
      const double w2 = 1.0 - w1;
      const double th0 = (double)a.t[c0] * pow(kP00 / p0, rovcp);
      const double th1 = (double)a.t[c1] * pow(kP00 / p1, rovcp);
      v = th0 * w1 + th1 * w2;
    } else {
      v = (double)a.t[cz] * pow(kP00 / (double)a.p[cz], rovcp);
    }
    a.th700[c2] = v;
  }
}

// ---------------------------------------------------------------------------
// Tail of the step.
template <typename R>
__device__ __forceinline__ R tv_factor(const QPtrs<R> &q, long long c,
                                       int ipptls) {
  const R one = R(1), ep1 = R(kEp1);
  if (ipptls > 1)
    return one + ep1 * q.q[0][c] - q.q[1][c] - q.q[2][c] - q.q[3][c] -
           q.q[4][c];
  if (ipptls > 0) return one + ep1 * q.q[0][c] - q.q[1][c];
  return one + ep1 * q.q[0][c];
}

// pfwsat: Share/pfwsat.inc (double inside, as the reference's rk8 locals)
__device__ __forceinline__ double pfwsat(double t, double p) {
  const double t_limit = t - kTzero;
  const double td = rmin(rmax(t_limit, -75.0), 100.0);
  double es;
  if (td >= 0.0)
    es = rmin(0.611213476e+03 + td * (0.444007856e+02 + td * (0.143064234e+01 +
              td * (0.264461437e-01 + td * (0.305903558e-03 + td * (0.196237241e-05 +
              td * (0.892344772e-08 + td * (-0.373208410e-10 + td * 0.209339997e-13))))))),
              0.15 * p);
  else
    es = rmin(0.611123516e+03 + td * (0.503109514e+02 + td * (0.188369801e+01 +
              td * (0.420547422e-01 + td * (0.614396778e-03 + td * (0.602780717e-05 +
              td * (0.387940929e-07 + td * (0.149436277e-09 + td * 0.262655803e-12))))))),
              0.15 * p);
  return kEp2 * (es / (p - es));
}

// flags: 1 = dycore tail (:1121-1125), 2 = diagnostics p, rho, qsat
// (:348-352), 4 = status update (:1410-1438).  Pointwise on cross points.
// p comes from the double Exner function in both forms.
// flags = 16: the three in one pass for a physics-off step, whose
// tendencies are zero: status_update's additions x + dtinc*0 leave t and q
// at the dycore tail's values and ux, vx at the advection's, so its
// uvxtouvstag would rewrite the u, v the advection's last back-staggering
// wrote; what remains is the species clamp and Tv, theta_v, rho, qsat from
// the clamped species.  No tendency array is read.  The surface pressure,
// which needs the dycore tail's Tv, runs before this pass (SurfacePressure
// with from_exner).
template <typename R>
struct Tail {
  Dims d;
  QPtrs<R> q;
  int nqx;
  int ipptls;
  R * t;
  R * tvirt;
  double * tetav;
  const double * pai;
  R * p;
  R * rho;
  R * qsat;
  R * ux;
  R * vx;
  const R * tten;
  const R * uten;
  const R * vten;
  double dtinc_;
  int flags;
  static __device__ __forceinline__ void run(const Tail &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void Tail<R>::run(const Tail &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const QPtrs<R> &q = args_.q;
  const int nqx = args_.nqx;
  const int ipptls = args_.ipptls;
  R *__restrict__ t = args_.t;
  R *__restrict__ tvirt = args_.tvirt;
  double *__restrict__ tetav = args_.tetav;
  const double *__restrict__ pai = args_.pai;
  R *__restrict__ p = args_.p;
  R *__restrict__ rho = args_.rho;
  R *__restrict__ qsat = args_.qsat;
  R *__restrict__ ux = args_.ux;
  R *__restrict__ vx = args_.vx;
  const R *__restrict__ tten = args_.tten;
  const R *__restrict__ uten = args_.uten;
  const R *__restrict__ vten = args_.vten;
  const double dtinc_ = args_.dtinc_;
  const int flags = args_.flags;

  const int j = 1 + bx * TJ + threadIdx.x;
  const int i = 1 + by * TI + threadIdx.y;
  const int k = 1 + bz;
  if (j > d.je2 || i > d.ie2) return;
  const R dtinc = R(dtinc_);
  const long long c = d.at(j, i, k);
  const bool interior = j >= d.jc1 && j <= d.jc2 && i >= d.ic1 && i <= d.ic2;
  // qxcheckval / qxzeroval: Main/mpplib/mod_runparams.F90:184-192
  const double qcheck[10] = {1.0e-8, 1.0e-16, 1.0e-16, 1.0e-16, 1.0e-16,
                             1.0e-16, 1.0e-16, 1.0e10, 100.0, 0.01};
  const double qzero[10] = {1.0e-8, 0.0, 0.0, 0.0, 0.0,
                            0.0, 0.0, 1.0e10, 100.0, 0.01};
  const double pa = pai[c];
  if (flags & 16) {
    // theta_v -> T -> Tv -> theta_v in double in both forms: T and Tv are
    // stored in the bulk type, but theta_v never passes through them (in
    // the mixed form a float T here re-rounded theta_v every step, which
    // was that form's largest error source).  With R = double these are
    // the reference's operations.
    // the species loops run over the fixed maximum with a guard, so that
    // qv and the check tables are indexed by constants and stay in
    // registers (a runtime bound made the compiler place them in LDS)
    R qv[10];
#pragma unroll
    for (int n = 0; n < 10; ++n)
      if (n < nqx) qv[n] = q.q[n][c];
    auto factor = [&]() -> double {
      const double one = 1.0, ep1 = kEp1;
      if (ipptls > 1)
        return one + ep1 * (double)qv[0] - (double)qv[1] - (double)qv[2] - (double)qv[3] -
               (double)qv[4];
      if (ipptls > 0) return one + ep1 * (double)qv[0] - (double)qv[1];
      return one + ep1 * (double)qv[0];
    };
    double tt;
    const R pp = R(pow(pa, kCpovr) * kP00);
    p[c] = pp;
    if (interior) {
      tt = (tetav[c] * pa) / factor();
      tt = tt + (double)dtinc * 0.0;
      t[c] = R(tt);
#pragma unroll
      for (int n = 0; n < 10; ++n) {
        if (n < nqx) {
          R a = qv[n] + dtinc * R(0);
          if (a < R(qcheck[n])) a = R(qzero[n]);
          if (a != qv[n] || signbit(a) != signbit(qv[n])) q.q[n][c] = a;
          qv[n] = a;
        }
      }
    } else {
      tt = (double)t[c];
    }
    const double tv = tt * factor();
    tvirt[c] = R(tv);
    tetav[c] = tv / pa;
    const R ts = R(tt);
    rho[c] = pp / (R(kRgas) * ts);
    qsat[c] = R(pfwsat((double)ts, (double)pp));
    return;
  }
  R tt = t[c];
  if (flags & 1) {
    const R tv = R(tetav[c] * pa);
    tvirt[c] = tv;
    if (interior) {
      tt = tv / tv_factor(q, c, ipptls);
      t[c] = tt;
    }
  }
  R pp = R(0);
  if (flags & 2) {
    pp = R(pow(pa, kCpovr) * kP00);
    p[c] = pp;
    rho[c] = pp / (R(kRgas) * tt);
    qsat[c] = R(pfwsat((double)tt, (double)pp));
  }
  if (flags & 32) {
    // restart (Main/mod_init.F90:951-955): tvirt from water vapour alone,
    // whatever ipptls, then theta_v on every cross point
    const R tv = tt * (R(1) + R(kEp1) * q.q[0][c]);
    tvirt[c] = tv;
    tetav[c] = (double)tv / pa;
  }
  if (flags & 8) {
    // temp_to_tvirt (:1608-1627) and tetav = tvirt/pai on every cross point
    const R tv = tt * tv_factor(q, c, ipptls);
    tvirt[c] = tv;
    tetav[c] = (double)tv / pa;
  }
  if (flags & 4) {
    if (!(flags & 2)) pp = p[c];
    if (interior) {
      tt = tt + dtinc * tten[c];
      t[c] = tt;
      ux[c] = ux[c] + dtinc * uten[c];
      vx[c] = vx[c] + dtinc * vten[c];
#pragma unroll
      for (int n = 0; n < 10; ++n) {
        if (n < nqx) {
          R a = q.q[n][c] + dtinc * q.qt[n][c];
          if (a < R(qcheck[n])) a = R(qzero[n]);
          q.q[n][c] = a;
        }
      }
    }
    const R tv = tt * tv_factor(q, c, ipptls);
    tvirt[c] = tv;
    tetav[c] = (double)tv / pa;
    rho[c] = pp / (R(kRgas) * tt);
    qsat[c] = R(pfwsat((double)tt, (double)pp));
  }
}

// extrapolate_surface_pressure (:1592-1606) on interior columns.  With
// from_exner, Tv(kz-1), Tv(kz) and p(kz) are formed from theta_v and the
// Exner function as the dycore tail and the diagnostics form them (the
// physics-off step runs this before its one-pass tail).
template <typename R>
struct SurfacePressure {
  Dims d;
  const R * tvirt;
  const R * z;
  const R * p;
  R * ps;
  const double * tetav;
  const double * pai;
  int from_exner;
  static __device__ __forceinline__ void run(const SurfacePressure &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void SurfacePressure<R>::run(const SurfacePressure &args_, int bx, int by, int bz) {
  const Dims d = args_.d;
  const R *__restrict__ tvirt = args_.tvirt;
  const R *__restrict__ z = args_.z;
  const R *__restrict__ p = args_.p;
  R *__restrict__ ps = args_.ps;

  const int j = d.jc1 + bx * TJ + threadIdx.x;
  const int i = d.ic1 + by * TI + threadIdx.y;
  if (j > d.jc2 || i > d.ic2) return;
  const int kz = d.kz;
  const long long ck = d.at(j, i, kz), cm = ck - d.plane;
  R tvm, tvk, pk;
  if (args_.from_exner) {
    const double pak = args_.pai[ck];
    tvm = R(args_.tetav[cm] * args_.pai[cm]);
    tvk = R(args_.tetav[ck] * pak);
    pk = R(pow(pak, kCpovr) * kP00);
  } else {
    tvm = tvirt[cm];
    tvk = tvirt[ck];
    pk = p[ck];
  }
  R lrt = (tvm - tvk) / (z[cm] - z[ck]);
  if (lrt < -R(kGovcp))
    lrt = -R(kGovcp);
  else if (lrt > R(-0.005))
    lrt = R(0.65) * lrt - R(0.35) * R(kLrate);
  const R tv = tvk - lrt * R(0.5) * z[ck];
  ps[d.at2(j, i)] = pk * exp(R(kGovr) * z[ck] / tv);
}

// TKE between its full levels and the half levels it is advected on
// (ibltyp 2): mode 0 is zstagtoh(tke, tkex) (Main/mod_moloch.F90:1445-1459)
// at the start of advection, mode 1 htozstag(tkex, tke) (:1461-1475) at
// its end; cross points jce x ice, one thread per point and level.
template <typename R>
struct TkeStagger {
  Dims d;
  R *tke;              // kz + 1 levels
  R *tkex;             // kz levels
  int mode;
  static __device__ __forceinline__ void run(const TkeStagger &a, int bx, int by, int bz) {
    const Dims d = a.d;
    const int j = 1 + bx * TJ + (int)threadIdx.x;
    const int i = 1 + by * TI + (int)threadIdx.y;
    const int k = 1 + bz, kz = d.kz;
    if (j > d.je2 || i > d.ie2) return;
    const long long P = d.plane, c = d.at(j, i, k);
    if (a.mode == 0) {
      const R *fl = a.tke;
      R hl;
      if (k == 1)
        hl = R(0.5) * (fl[c + P] + fl[c]);
      else if (k == kz)
        hl = R(0.5) * (fl[c + P] + fl[c]);
      else
        hl = R(0.5625) * (fl[c + P] + fl[c]) - R(0.0625) * (fl[c + 2 * P] + fl[c - P]);
      a.tkex[c] = hl;
    } else {
      const R *hl = a.tkex;
      if (k == 1) return;
      R fl;
      if (k == 2)
        fl = R(0.5) * (hl[c] + hl[c - P]);
      else if (k == kz)
        fl = R(0.5) * (hl[c] + hl[c - P]);
      else
        fl = R(0.5625) * (hl[c] + hl[c - P]) - R(0.0625) * (hl[c + P] + hl[c - 2 * P]);
      a.tke[c] = fl;
    }
  }
};

// The TKE and tracer lines of status_update (Main/mod_moloch.F90:1419-1430):
// on jci x ici, tke (levels 1 .. kz+1) += dtinc * tketen with the tkemin
// floor, each tracer (levels 1 .. kz) += dtinc * chiten with the zero floor.
template <typename R>
struct TracerUpdate {
  Dims d;
  R *tke;
  const R *tketen;
  R *tr[kMaxTr];
  const R *chiten;     // ntr blocks of kz levels
  int ntr;
  double dtinc, tkemin;
  static __device__ __forceinline__ void run(const TracerUpdate &a, int bx, int by, int bz) {
    const Dims d = a.d;
    const int j = 1 + bx * TJ + (int)threadIdx.x;
    const int i = 1 + by * TI + (int)threadIdx.y;
    const int k = 1 + bz, kz = d.kz;
    if (j < d.jc1 || j > d.jc2 || i < d.ic1 || i > d.ic2) return;
    const long long c = d.at(j, i, k);
    const R dtinc = R(a.dtinc);
    if (a.tke) {
      R x = a.tke[c] + dtinc * a.tketen[c];
      if (x < R(a.tkemin)) x = R(a.tkemin);
      a.tke[c] = x;
    }
    if (k > kz) return;
    const long long n3 = d.plane * kz;
    for (int n = 0; n < a.ntr; ++n) {
      R x = a.tr[n][c] + dtinc * a.chiten[n * n3 + c];
      if (x < R(0)) x = R(0);
      a.tr[n][c] = x;
    }
  }
};

template <typename R>
struct Fill {
  R * p;
  long long n;
  double value;
  long long nblk;   // blocks of the launch (grid-stride over its own blocks)
  static __device__ __forceinline__ void run(const Fill &args_, int bx, int by, int bz);
};

template <typename R>
__device__ __forceinline__ void Fill<R>::run(const Fill &args_, int bx, int by, int bz) {
  R *__restrict__ p = args_.p;
  const long long n = args_.n;
  const double value = args_.value;

  const long long idx = (long long)bx * blockDim.x + threadIdx.x;
  const long long stride = args_.nblk * blockDim.x;
  for (long long m = idx; m < n; m += stride) p[m] = R(value);
}

// ---------------------------------------------------------------------------
// Error-free transformations for the compensated budget sums.
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__device__ __forceinline__ DD dd_add(DD x, DD y) {
  DD s = two_sum(x.hi, y.hi);
  const double lo = s.lo + x.lo + y.lo;
  return two_sum(s.hi, lo);
}

// massck terms (Main/mod_massck.F90:77-160): per cross point, the dry and
// water mass of interior points and the boundary-flux terms of the edge
// points, each product evaluated in double as the reference writes it;
// summed as double-double per block into partial[block][4].
template <typename R>
__global__ __launch_bounds__(256) void k_massck(
    Dims d, QPtrs<R> q, int nqx, const R *__restrict__ u,
    const R *__restrict__ v, const R *__restrict__ dz,
    const R *__restrict__ rho, double dx, double dt,
    double *__restrict__ partial) {
  __shared__ DD sh[4][256];
  const int j = 1 + blockIdx.x * TJ + threadIdx.x;
  const int i = 1 + blockIdx.y * TI + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int tid = threadIdx.y * TJ + threadIdx.x;
  const int jx = d.jx, iy = d.iy;
  DD acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  if (j <= d.je2 && i <= d.ie2) {
    const long long c = d.at(j, i, k);
    const double dxsq = dx * dx;
    const double dzc = dz[c], rhoc = rho[c];
    const bool jin = j >= d.jc1 && j <= d.jc2, iin = i >= d.ic1 && i <= d.ic2;
    // boundary input only through physical edges (has_bdy*): none when periodic
    const bool edges = !d.crm;
    if (jin && iin) {
      acc[0] = two_sum(0.0, dxsq * dzc * rhoc);
      DD qs = {0, 0};
      for (int n = 0; n < nqx; ++n)
        qs = dd_add(qs, two_sum(0.0, (double)q.q[n][c] * dxsq * dzc * rhoc));
      acc[1] = qs;
    }
    DD da = {0, 0}, qa = {0, 0};
    if (edges && j == 1) {
      const double ub = u[d.at(1, i, k)];
      da = dd_add(da, two_sum(0.0, ub * dt * dx * dzc * rhoc));
      for (int n = 0; n < nqx; ++n)
        qa = dd_add(qa, two_sum(0.0, (double)q.q[n][c] * ub * dt * dx * dzc * rhoc));
    }
    if (edges && j == jx - 1) {
      const double ub = u[d.at(jx, i, k)];
      da = dd_add(da, two_sum(0.0, -(ub * dt * dx * dzc * rhoc)));
      for (int n = 0; n < nqx; ++n)
        qa = dd_add(qa, two_sum(0.0, -((double)q.q[n][c] * ub * dt * dx * dzc * rhoc)));
    }
    if (edges && i == 1 && jin) {
      const double vb = v[d.at(j, 1, k)];
      da = dd_add(da, two_sum(0.0, vb * dt * dx * dzc * rhoc));
      for (int n = 0; n < nqx; ++n)
        qa = dd_add(qa, two_sum(0.0, (double)q.q[n][c] * vb * dt * dx * dzc * rhoc));
    }
    if (edges && i == iy - 1 && jin) {
      const double vb = v[d.at(j, iy, k)];
      da = dd_add(da, two_sum(0.0, -(vb * dt * dx * dzc * rhoc)));
      for (int n = 0; n < nqx; ++n)
        qa = dd_add(qa, two_sum(0.0, -((double)q.q[n][c] * vb * dt * dx * dzc * rhoc)));
    }
    acc[2] = da;
    acc[3] = qa;
  }
  for (int m = 0; m < 4; ++m) sh[m][tid] = acc[m];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st)
      for (int m = 0; m < 4; ++m) sh[m][tid] = dd_add(sh[m][tid], sh[m][tid + st]);
    __syncthreads();
  }
  if (tid == 0) {
    const long long b = ((long long)blockIdx.z * gridDim.y + blockIdx.y) *
                            gridDim.x + blockIdx.x;
    for (int m = 0; m < 4; ++m) {
      partial[(b * 4 + m) * 2] = sh[m][0].hi;
      partial[(b * 4 + m) * 2 + 1] = sh[m][0].lo;
    }
  }
}

// max/min of ps over interior points, with a NaN/Inf count (:407-422).
template <typename R>
__global__ __launch_bounds__(256) void k_finite(Dims d,
                                                const R *__restrict__ ps,
                                                double *__restrict__ partial) {
  __shared__ double smax[256], smin[256], sbad[256];
  const int j = d.jc1 + blockIdx.x * TJ + threadIdx.x;
  const int i = d.ic1 + blockIdx.y * TI + threadIdx.y;
  const int tid = threadIdx.y * TJ + threadIdx.x;
  double mx = -HUGE_VAL, mn = HUGE_VAL, bad = 0.0;
  if (j <= d.jc2 && i <= d.ic2) {
    const double x = ps[d.at2(j, i)];
    if (!isfinite(x)) bad = 1.0;
    else { mx = x; mn = x; }
  }
  smax[tid] = mx; smin[tid] = mn; sbad[tid] = bad;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      smax[tid] = rmax(smax[tid], smax[tid + st]);
      smin[tid] = rmin(smin[tid], smin[tid + st]);
      sbad[tid] = sbad[tid] + sbad[tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    partial[b * 3] = smax[0];
    partial[b * 3 + 1] = smin[0];
    partial[b * 3 + 2] = sbad[0];
  }
}

// ============================================================ launchers ===
namespace {

dim3 grid_cross(const Dims &d, int nz) {
  return dim3((unsigned)((d.je2 + TJ - 1) / TJ),
              (unsigned)((d.ie2 + TI - 1) / TI), (unsigned)nz);
}
dim3 grid_from2(const Dims &d, int jhi, int ihi, int nz) {
  // tiles starting at the interior's first point (jc1, ic1) up to jhi, ihi
  // inclusive
  return dim3((unsigned)((jhi - d.jc1 + 1 + TJ - 1) / TJ),
              (unsigned)((ihi - d.ic1 + 1 + TI - 1) / TI), (unsigned)nz);
}
// blocks of cj interior columns along j, one interior row each
dim3 grid_columns(const Dims &d, int cj) {
  return dim3((unsigned)((d.jc2 - d.jc1 + 1 + cj - 1) / cj), (unsigned)(d.ic2 - d.ic1 + 1), 1);
}
const dim3 kBlock(TJ, TI, 1);

const double *vec(const mlc_handle *h, int which) {
  // vec1d layout: gzitak(kz+1) | gzitakh(kz) | xkdamp(kz) | xknu(kz) | ffilt(kz)
  const int kz = h->d.kz;
  const long long off[5] = {0, kz + 1, 2LL * kz + 1, 3LL * kz + 1, 4LL * kz + 1};
  return h->vec1d + off[which];
}

template <typename R>
QPtrs<R> qptrs(mlc_handle *h) {
  QPtrs<R> q{};
  for (int n = 0; n < h->g.nqx; ++n) {
    q.q[n] = h->cur_adv(A_QX0 + n).as<R>();
    q.qt[n] = h->qxten.as<R>() + (long long)n * h->d.plane * h->d.kz;
  }
  return q;
}

// Levels per thread of the level-chunked kernels: about `target`, evened out.
void chunks(const Dims &d, int target, int *nch, int *kc) {
  *nch = (d.kz + target - 1) / target;
  *kc = (d.kz + *nch - 1) / *nch;
}

}  // namespace

#define MLC_LAUNCH(slot, kern, grid, block, ...)                                  \
  do {                                                                            \
    hipEvent_t ea_, eb_;                                                          \
    prof_events(h, slot, &ea_, &eb_);                                             \
    if (ea_)                                                                      \
      hipExtLaunchKernelGGL(kern, grid, block, 0, h->stream, ea_, eb_, 0u, __VA_ARGS__); \
    else                                                                          \
      hipLaunchKernelGGL(kern, grid, block, 0, h->stream, __VA_ARGS__);           \
    MLC_HIP(hipGetLastError());                                                   \
    prof_push(h, slot, ea_, eb_);                                                 \
  } while (0)

template <typename T, typename S, int G> struct XcdOrder<WafVertical<T, S, G>> {
  static constexpr bool value = false;
};
template <typename S, int GF, int GD> struct XcdOrder<WafVerticalMixed<S, GF, GD>> {
  static constexpr bool value = false;
};
template <typename R> struct XcdOrder<Tail<R>> { static constexpr bool value = false; };
template <typename R> struct XcdOrder<Fill<R>> { static constexpr bool value = false; };
template <typename R> struct XcdOrder<SurfacePressure<R>> { static constexpr bool value = false; };
template <typename R> struct XcdOrder<BdyVal<R>> { static constexpr bool value = false; };
template <typename R> struct XcdOrder<SpecNudge<R>> { static constexpr bool value = false; };
template <typename R> struct XcdOrder<Slice<R>> { static constexpr bool value = false; };
// Slice<double> takes 152 VGPRs (3 waves per SIMD) unbounded; capped at 128
// it spills 22 to scratch and still runs a 256 x 256 x 41 domain in one
// round instead of 1.33 (170 -> 126 us).  Tail<float> reaches 6 without
// spilling.
template <> struct MinWaves<Slice<double>> { static constexpr int value = 4; };
template <> struct MinWaves<Tail<float>> { static constexpr int value = 6; };

// Launch (or, while the handle records, append to its launch record) one
// stage on the handle's stream.
template <class A>
int emit(mlc_handle *h, int slot, dim3 grid, dim3 block, size_t lds, const A &a) {
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return MLC_OK;
  if (h->rec) {
    LaunchRec r;
    r.slot = slot;
    r.many = reinterpret_cast<const void *>(&k_many<A>);
    r.grid = grid;
    r.block = block;
    r.lds = lds;
    r.args.resize(sizeof(A));
    std::memcpy(r.args.data(), &a, sizeof(A));
    h->rec->push_back(std::move(r));
    return MLC_OK;
  }
  hipEvent_t ea, eb;
  prof_events(h, slot, &ea, &eb);
  void (*kern)(A) = h->tune.xcd_map && XcdOrder<A>::value &&
                            (unsigned long long)grid.x * grid.y * grid.z >= 16
                        ? k_one_xcd<A>
                        : k_one<A>;
  if (ea)
    hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)lds, h->stream, ea, eb, 0u, a);
  else
    hipLaunchKernelGGL(kern, grid, block, lds, h->stream, a);
  MLC_HIP(hipGetLastError());
  prof_push(h, slot, ea, eb);
  return MLC_OK;
}

#define MLC_EMIT(slot, A, grid, block, lds, ...)                        \
  do {                                                                  \
    const int rc_ = emit(h, slot, grid, block, lds, A{__VA_ARGS__});    \
    if (rc_) return rc_;                                                \
  } while (0)

// The boundary-weight lookup of weight w (0 u, 1 v, 2 w) built by mlc_init.
template <typename R>
static BdyWt<R> bdy_wt(const mlc_handle *h, int w) {
  return BdyWt<R>{h->bidx + (long long)w * h->d.plane, h->btab[w].as<R>(), h->d.kz};
}

// The terrain metrics as the kernels form them: on for the stage `bit` of
// MLC_VAR_TERRAIN's mask (1 divergence, 2 column solve, 4 advection
// staggering, 8 horizontal WAF), which is non-zero only once verified.
static ZMet zmet(const mlc_handle *h, int bit) {
  const long long P = h->d.plane;
  return ZMet{h->ztab, h->ztab + 2 * h->d.kz + 1, h->hgeo, h->hgeo + P, h->hgeo + 2 * P,
              h->d.kz, (h->tune.terrain & bit) ? 1 : 0};
}

template <typename R>
static int sound_substep(mlc_handle *h, bool first) {
  const Dims d = h->d;
  const double dts = h->dtsound;
  const double dtrdx = dts * h->rdx;
  const double dtrdy = dts * h->rdx;
  const double dtrdz = dts * h->rdzita;
  const double zcs2 = (dtrdz * dtrdz) * kRdrcv;
  const double dxrdt = h->g.dx / dts;
  const int rot = h->g.lrotllr;
  double *tetav = h->cur_adv(A_TETAV).as<double>();
  double *pai = h->cur_adv(A_PAI).as<double>();
  int nch, kc;
  chunks(d, h->tune.level_chunk, &nch, &kc);
  MLC_EMIT(K_DIV, SoundDiv<R>, grid_cross(d, nch), kBlock, 0, d,
             h->u.as<R>(), h->v.as<R>(), h->rfmzu.as<R>(), h->rfmzv.as<R>(),
             h->fmz.as<R>(), h->mx.as<R>(), h->mx2.as<R>(), h->rmu.as<R>(),
             h->rmv.as<R>(), h->hx.as<R>(), h->hy.as<R>(), vec(h, 0),
             h->zdiv2.as<R>(), h->s.as<R>(), h->w, dtrdx, dtrdy, rot, kc, zmet(h, 1));
  // The requested column variant, or the nearest one whose LDS fits the
  // column height (all variants give identical results): the chunked solve
  // holds kz <= CK * CC levels, the three-phase one 4 (kz + 1) doubles per
  // column within 64 KB, the streamed one any height.
  int column = h->tune.column, column_j = h->tune.column_j;
  // 64 columns (a whole wave on the recurrences) hold up to 44 levels in
  // 11-level chunks, and their LDS (3 (kz - 1) 64 doubles) fits 64 KB up to 42
  if (column == 2 && column_j == 64 && d.kz > 42) column_j = 32;
  if (column == 2 && column_j == 32 && d.kz > kColK * kColC) column_j = 16;
  if (column == 0 && 4 * (size_t)(d.kz + 1) * kColJ * sizeof(double) > 65536) {
    column = 2;
    column_j = 16;
  }
  if (column == 2 && d.kz > 16 * kColC) column = 1;
  if (column == 2 && column_j == 64) {
    const dim3 cgrid = grid_columns(d, 64);
    const size_t lds = col_chunk_lds(d.kz, 64);
    using ColC = SoundColumnChunk<R, 64, 4, 11>;
    MLC_EMIT(K_COL, ColC, cgrid, dim3(64, 4, 1), lds, d,
             h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
             h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, dts, dtrdz,
             zcs2, h->g.do_divfilter, first ? 1 : 0, zmet(h, 2));
  } else if (column == 2 && column_j == 16) {
    const dim3 cgrid = grid_columns(d, 16);
    const size_t lds = col_chunk_lds(d.kz, 16);
    if (d.kz <= 16 * 3) {
      using ColC = SoundColumnChunk<R, 16, 16, 3>;
      MLC_EMIT(K_COL, ColC, cgrid, dim3(16, 16, 1), lds, d,
               h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
               h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, dts, dtrdz,
               zcs2, h->g.do_divfilter, first ? 1 : 0, zmet(h, 2));
    } else {
      using ColC = SoundColumnChunk<R, 16, 16>;
      MLC_EMIT(K_COL, ColC, cgrid, dim3(16, 16, 1), lds, d,
               h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
               h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, dts, dtrdz,
               zcs2, h->g.do_divfilter, first ? 1 : 0, zmet(h, 2));
    }
  } else if (column == 2) {
    const dim3 cgrid = grid_columns(d, kColJ);
    const size_t lds = col_chunk_lds(d.kz, kColJ);
    if (d.kz <= kColK * 6) {
      using ColC = SoundColumnChunk<R, kColJ, kColK, 6>;
      MLC_EMIT(K_COL, ColC, cgrid, dim3(kColJ, kColK, 1), lds, d,
               h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
               h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, dts, dtrdz,
               zcs2, h->g.do_divfilter, first ? 1 : 0, zmet(h, 2));
    } else {
      using ColC = SoundColumnChunk<R, kColJ, kColK>;
      MLC_EMIT(K_COL, ColC, cgrid, dim3(kColJ, kColK, 1), lds, d,
               h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
               h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, dts, dtrdz,
               zcs2, h->g.do_divfilter, first ? 1 : 0, zmet(h, 2));
    }
  } else if (column == 1) {
    MLC_EMIT(K_COL, SoundColumnStream<R>, grid_from2(d, d.jc2, d.ic2, 1), kBlock, 0,
             d, h->zdiv2.as<R>(), h->s.as<R>(), bdy_wt<R>(h, 2), (const double *)tetav,
             h->fmz64, h->fmzf64, vec(h, 3), vec(h, 4), h->tetavf, h->w, pai, h->colw,
             h->colw + h->d.plane * h->d.kz, h->colw + 2 * h->d.plane * h->d.kz, dts, dtrdz,
             zcs2, h->g.do_divfilter, first ? 1 : 0);
  } else {
    const dim3 cgrid = grid_columns(d, kColJ);
    const size_t lds = 4 * (size_t)(d.kz + 1) * kColJ * sizeof(double);
    MLC_EMIT(K_COL, SoundColumn<R>, cgrid, dim3(kColJ, kColK, 1), lds, d,
             h->zdiv2.as<R>(), h->s.as<R>(), (const double *)tetav, h->fmz64,
             h->fmzf64, bdy_wt<R>(h, 2), vec(h, 3), vec(h, 4), h->tetavf,
             h->w, pai, dts, dtrdz, zcs2, h->g.do_divfilter, first ? 1 : 0);
  }
  // the momentum update's own level chunk, if set: its best differs from the
  // divergence's (whole columns per thread in FP64 at 256x256x41)
  if (h->tune.mom_chunk > 0) chunks(d, h->tune.mom_chunk, &nch, &kc);
  MLC_EMIT(K_MOM, SoundMomentum<R>, grid_from2(d, d.je2, d.ie2, nch), kBlock, 0, d, h->u.as<R>(), h->v.as<R>(), h->zdiv2.as<R>(),
             (const double *)tetav, (const double *)pai, bdy_wt<R>(h, 0),
             bdy_wt<R>(h, 1), h->mu.as<R>(), h->mv.as<R>(), h->hx.as<R>(),
             h->hy.as<R>(), h->coru.as<R>(), h->corv.as<R>(), vec(h, 2),
             vec(h, 1), dts, dtrdx, dtrdy, dxrdt, h->g.do_divdamp, rot, kc);
  return MLC_OK;
}

// One sound() call (Main/mod_moloch.F90:545-736): mo_nsound acoustic
// sub-steps.
template <typename R>
static int sound_call(mlc_handle *h) {
  for (int n = 1; n <= h->g.nsound; ++n) {
    const int rc = sound_substep<R>(h, n == 1);
    if (rc) return rc;
  }
  return MLC_OK;
}

int launch_sound(mlc_handle *h) {
  return h->mixed ? sound_call<float>(h) : sound_call<double>(h);
}

template <typename R, int G>
static int waf_vertical(mlc_handle *h, const Dims &d, const AdvPtrs<R> &vb, int nb,
                        const AdvPtrs<double> &vp, int np, double dtrdz) {
  using WafVRR = WafVertical<R, R, G>;
  using WafVDR = WafVertical<double, R, G>;
  if constexpr (sizeof(R) == 4) {
    if (np && (h->tune.mixed_merge & 1)) {
      // float and double batches in one launch
      constexpr int GD = 2;
      using WafVM = WafVerticalMixed<R, G, GD>;
      const int ngf = (nb + G - 1) / G, ngd = (np + GD - 1) / GD;
      const WafVM a{WafVertical<float, R, G>{d, vb, h->s.as<R>(), h->zrfmu.as<R>(),
                                             h->zrfmd.as<R>(), dtrdz, nb},
                    WafVertical<double, R, GD>{d, vp, h->s.as<R>(), h->zrfmu.as<R>(),
                                               h->zrfmd.as<R>(), dtrdz, np},
                    ngf};
      return emit(h, K_WAF_V, grid_cross(d, ngf + ngd), kBlock, 0, a);
    }
  }
  MLC_EMIT(K_WAF_V, WafVRR, grid_cross(d, (nb + G - 1) / G), kBlock, 0, d, vb,
           h->s.as<R>(), h->zrfmu.as<R>(), h->zrfmd.as<R>(), dtrdz, nb);
  if (np)
    MLC_EMIT(K_WAF_V, WafVDR, grid_cross(d, (np + G - 1) / G), kBlock, 0, d, vp,
             h->s.as<R>(), h->zrfmu.as<R>(), h->zrfmd.as<R>(), dtrdz, np);
  return MLC_OK;
}

template <typename R, int TR, int PER>
static int waf_tile(mlc_handle *h, const Dims &d, const AdvPtrs<R> &hb,
                    const AdvPtrs<R> &wzb, int nb, const AdvPtrs<double> &hp,
                    const AdvPtrs<double> &wzp, int np, double dtrdx, double dtrdy,
                    int cux, int cvx, bool *curv) {
  using WafHRR = WafHorizontal<R, R, TR, PER>;
  using WafHDR = WafHorizontal<double, R, TR, PER>;
  const int rot = h->g.lrotllr;
  const unsigned gx = (unsigned)((d.je2 + TJ - 1) / TJ);
  const unsigned gy = (unsigned)((d.ie2 + TR - 1) / TR);
  // fields per block: all of the batch (0) or groups of waf_hgroup
  const int fg = h->tune.waf_hgroup > 0 ? h->tune.waf_hgroup : (nb > 0 ? nb : 1);
  const int ngb = nb > 0 ? (nb + fg - 1) / fg : 1;
  // the curvature pass fused into the ux, vx outputs when both fields are
  // in the same blocks (cux, cvx: their indices in the bulk batch, or -1)
  const bool fuse = cux >= 0 && cvx > cux && cux / fg == cvx / fg;
  *curv = fuse;
  const int fux = fuse ? cux : -1, fvx = fuse ? cvx : -1;
  const double dta = h->dtstepa;
  const R *tanrot = h->tanrot.as<R>();
  if constexpr (sizeof(R) == 4) {
    if (np && (h->tune.mixed_merge & 4) && h->tune.waf_hgroup == 0) {
      // the double batch in the float batch's blocks, one coefficient setup
      using WafHF = WafHorizontalFused<R, TR, PER>;
      const WafHF a{
          WafHorizontal<float, R, TR, PER>{d, hb, wzb, nb, 1, nb > 0 ? nb : 1, h->u.as<R>(),
                                      h->v.as<R>(), h->mu.as<R>(), h->mv.as<R>(),
                                      h->mx.as<R>(), h->mx2.as<R>(), h->rmu.as<R>(),
                                      h->rmv.as<R>(), h->rfmzu.as<R>(), h->rfmzv.as<R>(),
                                      h->fmz.as<R>(), dtrdx, dtrdy, rot, zmet(h, 8), fux, fvx,
                                      tanrot, h->rdx, dta},
          hp, wzp, np};
      return emit(h, K_WAF_H, dim3(gx, gy, (unsigned)d.kz), kBlock, 0, a);
    }
    if (np && (h->tune.mixed_merge & 2)) {
      // float and double batches in one launch
      const int fgp = h->tune.waf_hgroup > 0 ? h->tune.waf_hgroup : np;
      const int ngp = (np + fgp - 1) / fgp;
      using WafHM = WafHorizontalMixed<R, TR, PER>;
      const WafHM a{
          WafHorizontal<float, R, TR, PER>{d, hb, wzb, nb, ngb, fg, h->u.as<R>(), h->v.as<R>(),
                                      h->mu.as<R>(), h->mv.as<R>(), h->mx.as<R>(),
                                      h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(),
                                      h->rfmzu.as<R>(), h->rfmzv.as<R>(), h->fmz.as<R>(),
                                      dtrdx, dtrdy, rot, zmet(h, 8), fux, fvx, tanrot,
                                      h->rdx, dta},
          WafHorizontal<double, R, TR, PER>{d, hp, wzp, np, ngp, fgp, h->u.as<R>(), h->v.as<R>(),
                                       h->mu.as<R>(), h->mv.as<R>(), h->mx.as<R>(),
                                       h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(),
                                       h->rfmzu.as<R>(), h->rfmzv.as<R>(), h->fmz.as<R>(),
                                       dtrdx, dtrdy, rot, zmet(h, 8), -1, -1, tanrot, 0.0,
                                       0.0},
          (int)(gx * ngb)};
      return emit(h, K_WAF_H, dim3(gx * ngb + gx * ngp, gy, (unsigned)d.kz), kBlock, 0, a);
    }
  }
  MLC_EMIT(K_WAF_H, WafHRR, dim3(gx * ngb, gy, (unsigned)d.kz), kBlock, 0, d, hb, wzb, nb,
           ngb, fg, h->u.as<R>(), h->v.as<R>(), h->mu.as<R>(), h->mv.as<R>(), h->mx.as<R>(),
           h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(), h->rfmzu.as<R>(),
           h->rfmzv.as<R>(), h->fmz.as<R>(), dtrdx, dtrdy, rot, zmet(h, 8), fux, fvx,
           tanrot, h->rdx, dta);
  if (np) {
    const int fgp = h->tune.waf_hgroup > 0 ? h->tune.waf_hgroup : np;
    const int ngp = (np + fgp - 1) / fgp;
    MLC_EMIT(K_WAF_H, WafHDR, dim3(gx * ngp, gy, (unsigned)d.kz), kBlock, 0, d, hp, wzp, np,
             ngp, fgp, h->u.as<R>(), h->v.as<R>(), h->mu.as<R>(), h->mv.as<R>(),
             h->mx.as<R>(), h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(), h->rfmzu.as<R>(),
             h->rfmzv.as<R>(), h->fmz.as<R>(), dtrdx, dtrdy, rot, zmet(h, 8), -1, -1,
             tanrot, 0.0, 0.0);
  }
  return MLC_OK;
}

template <typename R, int G>
static int waf_rows(mlc_handle *h, const Dims &d, const AdvPtrs<R> &hb,
                    const AdvPtrs<R> &wzb, int nb, const AdvPtrs<double> &hp,
                    const AdvPtrs<double> &wzp, int np, double dtrdx, double dtrdy) {
  using WafRRR = WafRows<R, R, G>;
  using WafRDR = WafRows<double, R, G>;
  const int rot = h->g.lrotllr;
  const int rows = h->tune.waf_rows;
  const int ice2 = d.iy - 1, jce2 = d.jx - 1;
  const int nchunk = (ice2 + rows - 1) / rows;
  const unsigned gx = (unsigned)((jce2 + kWrOut - 1) / kWrOut);
  const unsigned gy = (unsigned)((nchunk + kWrWaves - 1) / kWrWaves);
  const dim3 rblk(64, kWrWaves, 1);
  const int ngb = (nb + G - 1) / G;
  MLC_EMIT(K_WAF_H, WafRRR, dim3(gx, gy, (unsigned)(d.kz * ngb)), rblk, 0, d, hb, wzb,
           nb, ngb, rows, h->u.as<R>(), h->v.as<R>(), h->mu.as<R>(), h->mv.as<R>(),
           h->mx.as<R>(), h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(),
           h->rfmzu.as<R>(), h->rfmzv.as<R>(), h->fmz.as<R>(), dtrdx, dtrdy, rot);
  if (np) {
    const int ngp = (np + G - 1) / G;
    MLC_EMIT(K_WAF_H, WafRDR, dim3(gx, gy, (unsigned)(d.kz * ngp)), rblk, 0, d, hp,
             wzp, np, ngp, rows, h->u.as<R>(), h->v.as<R>(), h->mu.as<R>(), h->mv.as<R>(),
             h->mx.as<R>(), h->mx2.as<R>(), h->rmu.as<R>(), h->rmv.as<R>(),
             h->rfmzu.as<R>(), h->rfmzv.as<R>(), h->fmz.as<R>(), dtrdx, dtrdy, rot);
  }
  return MLC_OK;
}

template <typename R>
static int advection(mlc_handle *h) {

  const Dims d = h->d;
  const double dta = h->dtstepa;
  const double dtrdx = dta * h->rdx;
  const double dtrdy = dta * h->rdx;
  double dtrdz = dta * h->rdzita;
  dtrdz = 0.5 * dtrdz;
  const int rot = h->g.lrotllr;
  int snch, skc;   // level chunks of the staggering passes
  chunks(d, h->tune.stag_chunk, &snch, &skc);
  MLC_EMIT(K_STAG_IN, StaggerIn<R>, grid_cross(d, snch), kBlock, 0, d,
             h->u.as<R>(), h->v.as<R>(), (const double *)h->w,
             h->fmzf.as<R>(), h->s.as<R>(), h->cur_adv(A_UX).as<R>(),
             h->cur_adv(A_VX).as<R>(), h->cur_adv(A_WX).as<R>(), 0, zmet(h, 4), skc);
  if (h->s_tkex >= 0)
    MLC_EMIT(K_STAG_IN, TkeStagger<R>, grid_cross(d, d.kz), kBlock, 0, d, h->tke.as<R>(),
             h->cur_adv(h->s_tkex).as<R>(), 0);
  // batches: the fields stored in the bulk type, and (mixed form) the
  // double fields tetav and pai, at most kMaxAdv fields per launch (the
  // fields are independent, so the split is exact)
  int bl[kMaxSlots], nbl = 0, dl[2], ndl = 0;
  for (int m = 0; m < h->nadvf; ++m) {
    const int f = h->adv_list[m];
    if (h->mixed && (f == A_PAI || f == A_TETAV))
      dl[ndl++] = f;
    else
      bl[nbl++] = f;
  }
  bool curv = false;          // the curvature pass ran inside the horizontal WAF
  for (int c0 = 0; c0 < nbl; c0 += kMaxAdv) {
    AdvPtrs<R> vb{}, hb{}, wzb{};
    AdvPtrs<double> vp{}, hp{}, wzp{};
    int nb = 0, np = 0, cux = -1, cvx = -1;
    for (int m = c0; m < nbl && m < c0 + kMaxAdv; ++m) {
      const int f = bl[m];
      if (f == A_UX) cux = nb;
      if (f == A_VX) cvx = nb;
      vb.src[nb] = h->cur_adv(f).as<R>();
      vb.dst[nb] = h->wz2[f].as<R>();
      hb.src[nb] = h->cur_adv(f).as<R>();
      hb.dst[nb] = h->alt_adv(f).as<R>();
      wzb.src[nb] = h->wz2[f].as<R>();
      ++nb;
    }
    for (int m = 0; c0 == 0 && m < ndl; ++m) {
      const int f = dl[m];
      vp.src[np] = h->cur_adv(f).as<double>();
      vp.dst[np] = h->wz2[f].as<double>();
      hp.src[np] = h->cur_adv(f).as<double>();
      hp.dst[np] = h->alt_adv(f).as<double>();
      wzp.src[np] = h->wz2[f].as<double>();
      ++np;
    }
    {
      int rc = h->tune.wafv_group == 1 ? waf_vertical<R, 1>(h, d, vb, nb, vp, np, dtrdz)
             : h->tune.wafv_group == 2 ? waf_vertical<R, 2>(h, d, vb, nb, vp, np, dtrdz)
                                       : waf_vertical<R, 4>(h, d, vb, nb, vp, np, dtrdz);
      if (rc) return rc;
    }
    if (h->tune.waf_h == 0 && !d.crm) {
      // row-streaming form (bounded domains only; a periodic domain takes
      // the tile form, whose halo loads wrap)
      int rc = h->tune.waf_group == 2
                   ? waf_rows<R, 2>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy)
                   : waf_rows<R, 4>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy);
      if (rc) return rc;
    } else {
      // LDS-tile form
      int rc;
      bool fused = false;
      // with the row-streaming form selected the curvature pass runs on its
      // own also on a periodic domain, so that every member of a batch
      // records the same launches
      if (h->tune.waf_h == 0) cux = cvx = -1;
      if (d.crm)
        rc = h->tune.waf_tile == 8
                 ? waf_tile<R, 8, 1>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy, cux, cvx, &fused)
                 : waf_tile<R, 4, 1>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy, cux, cvx, &fused);
      else
        rc = h->tune.waf_tile == 8
                 ? waf_tile<R, 8, 0>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy, cux, cvx, &fused)
                 : waf_tile<R, 4, 0>(h, d, hb, wzb, nb, hp, wzp, np, dtrdx, dtrdy, cux, cvx, &fused);
      if (rc) return rc;
      curv = curv || fused;
    }
  }
  for (int m = 0; m < h->nadvf; ++m) h->cur[h->adv_list[m]] ^= 1;
  if (!curv)
    MLC_EMIT(K_CURV, Curvature<R>, grid_from2(d, d.jc2, d.ic2, d.kz), kBlock, 0, d, h->cur_adv(A_UX).as<R>(), h->cur_adv(A_VX).as<R>(),
             h->mu.as<R>(), h->mv.as<R>(), h->tanrot.as<R>(), h->rdx, dta, rot);
  MLC_EMIT(K_STAG_OUT, StaggerOut<R>, grid_cross(d, snch), kBlock, 0, d,
             h->cur_adv(A_UX).as<R>(), h->cur_adv(A_VX).as<R>(),
             h->cur_adv(A_WX).as<R>(), h->u.as<R>(), h->v.as<R>(), h->w, 1, skc);
  if (h->s_tkex >= 0)
    MLC_EMIT(K_STAG_OUT, TkeStagger<R>, grid_cross(d, d.kz), kBlock, 0, d, h->tke.as<R>(),
             h->cur_adv(h->s_tkex).as<R>(), 1);
  return MLC_OK;
}

int launch_advection(mlc_handle *h) {
  return h->mixed ? advection<float>(h) : advection<double>(h);
}

template <typename R>
static int waf_metrics(mlc_handle *h) {
  const Dims d = h->d;
  double dtrdz = h->dtstepa * h->rdzita;
  dtrdz = 0.5 * dtrdz;
  MLC_EMIT(K_FILL, WafMetrics<R>, grid_cross(d, d.kz), kBlock, 0, d,
             (const double *)h->fmz64, (const double *)h->fmzf64,
             h->zrfmu.as<R>(), h->zrfmd.as<R>(), dtrdz);
  return MLC_OK;
}

int launch_waf_metrics(mlc_handle *h) {
  return h->mixed ? waf_metrics<float>(h) : waf_metrics<double>(h);
}

// mlc_init's check of the terrain metrics: every point of fmz (jce, ice, kz),
// fmzf (jce, ice, kz+1), rfmzu (jde, ice, kz) and rfmzv (jce, ide, kz), and of
// the double copies the mixed form's column solve reads, compared bit for
// bit with the value the kernels would form from the orography (in the
// arrays' type).  One 64 x 4 tile of one level per block; each block writes
// the number of its threads that found a difference to out[block].
template <typename R>
__global__ __launch_bounds__(256) void k_terrain_check(Dims d, ZMet zm, const R *fmz,
                                                       const R *fmzf, const R *rfmzu,
                                                       const R *rfmzv, const double *fmz64,
                                                       const double *fmzf64, double *out) {
  const int j = 1 + blockIdx.x * TJ + threadIdx.x;
  const int i = 1 + blockIdx.y * TI + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int kz = d.kz;
  bool bad = false;
  if (j <= d.jx && i <= d.iy) {
    const long long c = d.at(j, i, k), c2 = d.at2(j, i);
    auto differs = [](R stored, double formed) {
      if constexpr (sizeof(R) == 4)
        return __float_as_uint(stored) != __float_as_uint(float(formed));
      else
        return __double_as_longlong(stored) != __double_as_longlong(formed);
    };
    auto differs64 = [](double stored, double formed) {
      return __double_as_longlong(stored) != __double_as_longlong(formed);
    };
    const bool jc = j <= d.je2, ic = i <= d.ie2;
    if (jc && ic && k <= kz) {
      bad |= differs(fmz[c], zm.fmz(c2, k));
      if (fmz64) bad |= differs64(fmz64[c], zm.fmz(c2, k));
    }
    if (jc && ic) {
      bad |= differs(fmzf[c], zm.fmzf(c2, k));
      if (fmzf64) bad |= differs64(fmzf64[c], zm.fmzf(c2, k));
    }
    if (ic && k <= kz) bad |= differs(rfmzu[c], zm.rfmzu(c2, k));
    if (jc && k <= kz) bad |= differs(rfmzv[c], zm.rfmzv(c2, k));
  }
  const int n = __syncthreads_count(bad ? 1 : 0);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    out[((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = (double)n;
}

// Points where the formed terrain metrics differ from the uploaded ones
// (0: the kernels may form them).  Synchronous.
int launch_terrain_check(mlc_handle *h, long long *nbad) {
  const Dims d = h->d;
  ZMet zm = zmet(h, 0);
  zm.on = 1;
  const dim3 grid((unsigned)((d.jx + TJ - 1) / TJ), (unsigned)((d.iy + TI - 1) / TI),
                  (unsigned)(d.kz + 1));
  const size_t nblk = (size_t)grid.x * grid.y * grid.z;
  if (nblk > h->red_len) return fail(MLC_E_ARG, "terrain check: reduction buffer too small");
  if (h->mixed)
    hipLaunchKernelGGL(k_terrain_check<float>, grid, kBlock, 0, h->stream, d, zm,
                       h->fmz.as<float>(), h->fmzf.as<float>(), h->rfmzu.as<float>(),
                       h->rfmzv.as<float>(), (const double *)h->fmz64,
                       (const double *)h->fmzf64, h->red);
  else
    hipLaunchKernelGGL(k_terrain_check<double>, grid, kBlock, 0, h->stream, d, zm,
                       h->fmz.as<double>(), h->fmzf.as<double>(), h->rfmzu.as<double>(),
                       h->rfmzv.as<double>(), (const double *)nullptr,
                       (const double *)nullptr, h->red);
  MLC_HIP(hipGetLastError());
  std::vector<double> cnt(nblk);
  MLC_HIP(hipMemcpyAsync(cnt.data(), h->red, nblk * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  MLC_HIP(hipStreamSynchronize(h->stream));
  long long n = 0;
  for (double v : cnt) n += (long long)v;
  *nbad = n;
  return MLC_OK;
}

template <typename R>
static int tail(mlc_handle *h, double dtinc, int flags) {
  const Dims d = h->d;
  MLC_EMIT(K_TAIL, Tail<R>, grid_cross(d, d.kz), kBlock, 0, d, qptrs<R>(h),
             h->g.nqx, h->g.ipptls, h->t.as<R>(), h->tvirt.as<R>(),
             h->cur_adv(A_TETAV).as<double>(),
             (const double *)h->cur_adv(A_PAI).as<double>(), h->p.as<R>(),
             h->rho.as<R>(), h->qsat.as<R>(), h->cur_adv(A_UX).as<R>(),
             h->cur_adv(A_VX).as<R>(), h->tten.as<R>(), h->uten.as<R>(),
             h->vten.as<R>(), dtinc, flags);
  if ((flags & 2) && !(flags & 32))   // a restart keeps the saved ps
    MLC_EMIT(K_PS, SurfacePressure<R>, grid_from2(d, d.jc2, d.ic2, 1), kBlock, 0,
               d, h->tvirt.as<R>(), h->zeta.as<R>(), h->p.as<R>(),
               h->ps.as<R>(), nullptr, nullptr, 0);
  if (flags & 4) {
    int snch, skc;   // level chunks of the staggering passes
    chunks(d, h->tune.stag_chunk, &snch, &skc);
    MLC_EMIT(K_UVSTAG, StaggerOut<R>, grid_cross(d, snch), kBlock, 0, d,
               h->cur_adv(A_UX).as<R>(), h->cur_adv(A_VX).as<R>(),
               h->cur_adv(A_WX).as<R>(), h->u.as<R>(), h->v.as<R>(), h->w, 0, skc);
  }
  return MLC_OK;
}

template <typename R>
static int lbc(mlc_handle *h, double x1) {
  const Dims d = h->d;
  const LbcData bd = {h->lbc.ps0, h->lbc.ps1, h->lbc.u0, h->lbc.u1, h->lbc.v0, h->lbc.v1,
                      h->lbc.t0, h->lbc.t1, h->lbc.pai0, h->lbc.pai1, h->lbc.qv0,
                      h->lbc.qv1, h->lbc.qc0, h->lbc.qc1, h->lbc.qi0, h->lbc.qi1,
                      h->lbc.xfu, h->lbc.xfv, h->lbc.xfc, h->lbc.tnudge};
  const mlc_lbc_desc &L = h->lbc.desc;
  R *u = h->u.as<R>(), *v = h->v.as<R>(), *t = h->t.as<R>();
  double *pai = h->cur_adv(A_PAI).as<double>();
  const dim3 b256(256, 1, 1);
  for (int sn = 0; sn < 2; ++sn) {
    const int n_along = sn ? d.jx : d.iy;
    const dim3 g((unsigned)((n_along + 256) / 256), (unsigned)d.kz, 1);
    MLC_EMIT(K_LBC, BdyVal<R>, g, b256, 0, d, bd, u, v, t, pai, h->w, qptrs<R>(h),
             h->ps.as<R>(), x1, h->g.nqx, h->g.ipptls, L.present_qc, L.present_qi, sn,
             h->s_tkex >= 0 ? h->tke.as<R>() : nullptr, h->tkemin);
  }
  MLC_EMIT(K_LBC, BdyRelax<R>, grid_from2(d, d.je2, d.ie2, d.kz), kBlock, 0, d, bd, u,
           v, t, pai, h->w, qptrs<R>(h), x1, h->g.dtsec, h->g.ipptls, L.present_qc,
           L.present_qi, L.top_nudge, L.nztop);
  if (L.spectral_nudge) {                                       // :497-504
    h->lbc.tspectral = h->lbc.tspectral + h->g.dtsec;
    if ((int)std::fmod(h->lbc.tspectral, L.dtrad) == 0) {
      const int jx = d.jx, iy = d.iy;
      // (field, b0, b1, j1, j2, i1, i2, jj1, jj2, ii1, ii2); the theta call
      // sums over j = jci1..jci1 only, as the reference does
      struct Call { R *f; const double *b0, *b1; int r[8]; } calls[3] = {
          {t, h->lbc.t0, h->lbc.t1, {1, jx - 1, 1, iy - 1, 2, 2, 2, iy - 2}},
          {u, h->lbc.u0, h->lbc.u1, {1, jx, 1, iy - 1, 2, jx - 1, 2, iy - 2}},
          {v, h->lbc.v0, h->lbc.v1, {1, jx - 1, 1, iy, 2, jx - 2, 2, iy - 1}}};
      const int km2 = 2 * L.km, lm2 = 2 * L.lm;
      const unsigned gj = (unsigned)((jx + 255) / 256), gi = (unsigned)((iy + 255) / 256);
      for (auto &cl : calls) {
        const dim3 grids[5] = {dim3(gj, (unsigned)iy, (unsigned)d.kz),
                               dim3(gi, (unsigned)km2, (unsigned)d.kz),
                               dim3(gj, (unsigned)iy, (unsigned)d.kz),
                               dim3(gj, (unsigned)lm2, (unsigned)d.kz),
                               dim3(gj, (unsigned)iy, (unsigned)d.kz)};
        for (int st = 0; st < 5; ++st)
          MLC_EMIT(K_LBC, SpecNudge<R>, grids[st], b256, 0, d, cl.f, cl.b0, cl.b1,
                   (const double *)h->lbc.bvx, (const double *)h->lbc.bvy,
                   (const double *)h->lbc.cnudge, h->lbc.zn, h->lbc.sx, h->lbc.sy, x1,
                   cl.r[0], cl.r[1], cl.r[2], cl.r[3], cl.r[4], cl.r[5], cl.r[6], cl.r[7],
                   km2, lm2, st);
      }
    }
  }
  int snch, skc;   // level chunks of the staggering passes
  chunks(d, h->tune.stag_chunk, &snch, &skc);
  MLC_EMIT(K_LBC, StaggerIn<R>, grid_cross(d, snch), kBlock, 0, d, u, v, (const double *)h->w,
           h->fmzf.as<R>(), h->s.as<R>(), h->cur_adv(A_UX).as<R>(), h->cur_adv(A_VX).as<R>(),
           h->cur_adv(A_WX).as<R>(), 1, zmet(h, 4), skc);
  return tail<R>(h, 0.0, 8);
}

int launch_boundary(mlc_handle *h, double x1) {
  return h->mixed ? lbc<float>(h, x1) : lbc<double>(h, x1);
}

template <typename R>
static int mkslice(mlc_handle *h, const mlc_slice_desc &sd) {
  const Dims d = h->d;
  const dim3 grid((unsigned)((d.jx + 63) / 64), (unsigned)d.iy, 1);
  MLC_EMIT(K_SLICE, Slice<R>, grid, dim3(64, kSliceWaves, 1), 0, d, (const double *)h->cur_adv(A_PAI).as<double>(),
           (const double *)h->w, (const double *)h->sl.zetaf, (const double *)h->sl.xlat,
           (const R *)h->t.as<R>(), (const R *)h->p.as<R>(), (const R *)h->rho.as<R>(),
           (const R *)h->qsat.as<R>(), (const R *)h->zeta.as<R>(), (const R *)h->ps.as<R>(),
           qptrs<R>(h), h->sl.pf, h->sl.th, h->sl.rh, h->sl.omega, h->sl.rhox, h->sl.tp,
           h->sl.th700, h->sl.ptrop, h->sl.ktrop, h->sl.kmxpbl, sd.rhmin, sd.rhmax, sd.calday,
           sd.dayspy, h->g.nqx, sd.icldmstrat, sd.irceideal, sd.ibltyp);
  return MLC_OK;
}

int launch_mkslice(mlc_handle *h, const mlc_slice_desc &sd) {
  return h->mixed ? mkslice<float>(h, sd) : mkslice<double>(h, sd);
}

int launch_dyn_tail(mlc_handle *h) {
  return h->mixed ? tail<float>(h, 0.0, 1) : tail<double>(h, 0.0, 1);
}
int launch_diagnostics(mlc_handle *h) {
  return h->mixed ? tail<float>(h, 0.0, 2) : tail<double>(h, 0.0, 2);
}
// init's restart branch (Main/mod_init.F90:941-955): diagnostics, then Tv
// from water vapour alone and theta_v
int launch_init_restart(mlc_handle *h) {
  return h->mixed ? tail<float>(h, 0.0, 2 | 32) : tail<double>(h, 0.0, 2 | 32);
}
int launch_dyn_tail_diagnostics(mlc_handle *h) {
  return h->mixed ? tail<float>(h, 0.0, 1 | 2) : tail<double>(h, 0.0, 1 | 2);
}
template <typename R>
static int tracer_update(mlc_handle *h, double dtinc) {
  if (h->s_tkex < 0 && h->s_tr0 < 0) return MLC_OK;
  const Dims d = h->d;
  TracerUpdate<R> a{};
  a.d = d;
  a.tke = h->s_tkex >= 0 ? h->tke.as<R>() : nullptr;
  a.tketen = h->s_tkex >= 0 ? h->tketen.as<R>() : nullptr;
  a.ntr = h->s_tr0 >= 0 ? h->g.ntr : 0;
  for (int n = 0; n < a.ntr; ++n) a.tr[n] = h->cur_adv(h->s_tr0 + n).as<R>();
  a.chiten = a.ntr ? h->chiten.as<R>() : nullptr;
  a.dtinc = dtinc;
  a.tkemin = h->tkemin;
  return emit(h, K_TAIL, grid_cross(d, d.kz + (a.tke ? 1 : 0)), kBlock, 0, a);
}

// The tail of a physics-off step (zero tendencies): surface pressure from
// the dycore tail's Tv, then the dycore tail, diagnostics and status_update
// in one pass, then the TKE / tracer update.
template <typename R>
static int step_tail(mlc_handle *h) {
  const Dims d = h->d;
  const double *tetav = h->cur_adv(A_TETAV).as<double>();
  const double *pai = h->cur_adv(A_PAI).as<double>();
  MLC_EMIT(K_PS, SurfacePressure<R>, grid_from2(d, d.jc2, d.ic2, 1), kBlock, 0, d,
           nullptr, h->zeta.as<R>(), nullptr, h->ps.as<R>(), tetav, pai, 1);
  MLC_EMIT(K_TAIL, Tail<R>, grid_cross(d, d.kz), kBlock, 0, d, qptrs<R>(h), h->g.nqx,
           h->g.ipptls, h->t.as<R>(), h->tvirt.as<R>(), h->cur_adv(A_TETAV).as<double>(), pai,
           h->p.as<R>(), h->rho.as<R>(), h->qsat.as<R>(), nullptr, nullptr, nullptr, nullptr,
           nullptr, h->g.dtsec, 16);
  return tracer_update<R>(h, h->g.dtsec);
}

int launch_step_tail(mlc_handle *h) {
  return h->mixed ? step_tail<float>(h) : step_tail<double>(h);
}

int launch_status_update(mlc_handle *h, double dtinc) {
  const int rc = h->mixed ? tail<float>(h, dtinc, 4) : tail<double>(h, dtinc, 4);
  if (rc) return rc;
  return h->mixed ? tracer_update<float>(h, dtinc) : tracer_update<double>(h, dtinc);
}

template <typename R>
static int fill_t(mlc_handle *h, R *p, long long n, double value) {
  if (n <= 0) return MLC_OK;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  MLC_EMIT(K_FILL, Fill<R>, dim3((unsigned)blocks), dim3(256), 0, p, n, value, blocks);
  return MLC_OK;
}

int launch_fill(mlc_handle *h, const Buf &b, long long n, double value) {
  return b.es == 4 ? fill_t<float>(h, b.as<float>(), n, value)
                   : fill_t<double>(h, b.as<double>(), n, value);
}
int launch_fill64(mlc_handle *h, double *p, long long n, double value) {
  return fill_t<double>(h, p, n, value);
}

template <typename R>
static int massck(mlc_handle *h, double out[4]) {
  const Dims d = h->d;
  const dim3 grid = grid_cross(d, d.kz);
  const size_t nb = (size_t)grid.x * grid.y * grid.z;
  if (h->red_len < nb * 8) return fail(MLC_E_STATE, "reduction buffer too small");
  MLC_LAUNCH(K_MASSCK, k_massck<R>, grid, kBlock, d, qptrs<R>(h), h->g.nqx,
             h->u.as<R>(), h->v.as<R>(), h->dz.as<R>(), h->rho.as<R>(),
             h->g.dx, h->g.dtsec, h->red);
  std::vector<double> part(nb * 8);
  MLC_HIP(hipMemcpyAsync(part.data(), h->red, nb * 8 * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream));
  MLC_HIP(hipStreamSynchronize(h->stream));
  // combine the block partials as double-double in block order
  for (int m = 0; m < 4; ++m) {
    double hi = 0.0, lo = 0.0;
    for (size_t b = 0; b < nb; ++b) {
      const double a = part[(b * 4 + m) * 2], al = part[(b * 4 + m) * 2 + 1];
      const double s = hi + a;
      const double bb = s - hi;
      const double e = (hi - (s - bb)) + (a - bb);
      lo = lo + e + al;
      hi = s;
    }
    out[m] = hi + lo;
  }
  return MLC_OK;
}

int launch_massck(mlc_handle *h, double out[4]) {
  return h->mixed ? massck<float>(h, out) : massck<double>(h, out);
}

// ---------------------------------------------------------------------------
// Allocation-free budget (mlc_massck_noalloc).  The block partials of
// k_massck and the four combined totals live in `scratch`, a file-scope
// device array: one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress
// returns for the current device, looked up once per device.
constexpr long long kNoAllocMaxBlocks = 131072;   // 8 MiB of partials
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocMaxBlocks * 8 + 8];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocMaxDevices] = {};

static int scratch_address(int dev, double **out) {
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return fail(MLC_E_ARG, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MLC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                 hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return MLC_OK;
}

// Combines the nb block partials of each quantity in block order with the
// same two-sum the host combine of massck() uses, so both entries return
// the same bits.  Thread m handles quantity m.
__global__ __launch_bounds__(64) void k_massck_combine(const double *__restrict__ partial,
                                                       long long nb,
                                                       double *__restrict__ total) {
  const int m = threadIdx.x;
  if (m >= 4) return;
  double hi = 0.0, lo = 0.0;
  for (long long b = 0; b < nb; ++b) {
    const double a = partial[(b * 4 + m) * 2], al = partial[(b * 4 + m) * 2 + 1];
    const double s = hi + a;
    const double bb = s - hi;
    const double e = (hi - (s - bb)) + (a - bb);
    lo = lo + e + al;
    hi = s;
  }
  total[m] = hi + lo;
}

template <typename R>
static int massck_noalloc(mlc_handle *h, double out[4]) {
  const Dims d = h->d;
  const dim3 grid = grid_cross(d, d.kz);
  const long long nb = (long long)grid.x * grid.y * grid.z;
  if (nb > kNoAllocMaxBlocks)
    return fail(MLC_E_ARG, "domain too large for the allocation-free budget: " +
                               std::to_string(nb) + " blocks of partials, at most " +
                               std::to_string(kNoAllocMaxBlocks));
  int dev = -1;
  MLC_HIP(hipGetDevice(&dev));
  if (dev != h->device)
    return fail(MLC_E_ARG, "the current device is not the handle's device");
  double *ptr = nullptr;
  if (int rc = scratch_address(dev, &ptr)) return rc;
  double *total = ptr + kNoAllocMaxBlocks * 8;
  // plain launches: no profiling events are made here
  hipLaunchKernelGGL(k_massck<R>, grid, kBlock, 0, h->stream, d, qptrs<R>(h), h->g.nqx,
                     h->u.as<R>(), h->v.as<R>(), h->dz.as<R>(), h->rho.as<R>(),
                     h->g.dx, h->g.dtsec, ptr);
  MLC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_massck_combine, dim3(1), dim3(64), 0, h->stream,
                     (const double *)ptr, nb, total);
  MLC_HIP(hipGetLastError());
  MLC_HIP(hipMemcpyAsync(out, total, 4 * sizeof(double), hipMemcpyDeviceToHost,
                         h->stream));
  MLC_HIP(hipStreamSynchronize(h->stream));
  return MLC_OK;
}

int launch_massck_noalloc(mlc_handle *h, double out[4]) {
  return h->mixed ? massck_noalloc<float>(h, out) : massck_noalloc<double>(h, out);
}

template <typename R>
static int check_finite(mlc_handle *h, double *psmax, double *psmin) {
  const Dims d = h->d;
  const dim3 grid = grid_from2(d, d.jc2, d.ic2, 1);
  const size_t nb = (size_t)grid.x * grid.y;
  if (h->red_len < nb * 3) return fail(MLC_E_STATE, "reduction buffer too small");
  MLC_LAUNCH(K_FINITE, k_finite<R>, grid, kBlock, d, h->ps.as<R>(), h->red);
  std::vector<double> part(nb * 3);
  MLC_HIP(hipMemcpyAsync(part.data(), h->red, nb * 3 * sizeof(double),
                         hipMemcpyDeviceToHost, h->stream));
  MLC_HIP(hipStreamSynchronize(h->stream));
  double mx = -HUGE_VAL, mn = HUGE_VAL, bad = 0.0;
  for (size_t b = 0; b < nb; ++b) {
    mx = part[b * 3] > mx ? part[b * 3] : mx;
    mn = part[b * 3 + 1] < mn ? part[b * 3 + 1] : mn;
    bad += part[b * 3 + 2];
  }
  if (psmax) *psmax = bad > 0 ? NAN : mx;
  if (psmin) *psmin = bad > 0 ? NAN : mn;
  if (bad > 0) return fail(MLC_E_NONFINITE, "CFL VIOLATION: non-finite surface pressure");
  return MLC_OK;
}

int launch_check_finite(mlc_handle *h, double *psmax, double *psmin) {
  return h->mixed ? check_finite<float>(h, psmax, psmin)
                  : check_finite<double>(h, psmax, psmin);
}

}  // namespace mlc
