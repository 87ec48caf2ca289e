// Horizontal diffusion of the MM cores (Main/mod_diffusion.F90) on the
// device: initialize_diffusion (:99-163), calc_coeff (:166-248), diffu_d
// (:263-511) and the diffu_x family (:513-913), for one domain that owns all
// four lateral boundaries.  FP64, compiled without contraction, so every
// coefficient and tendency is the reference's value bit for bit.
//
// Layout as the step's: [nk][iy][pj] with the pitch pj = jx rounded up to
// 32 doubles; F / FTEN hold nspec species of kz+1 levels each.
//
// Blocks of 64 x 4 threads along (j, i) and one grid z-slice per level (and
// species), so every wavefront reads contiguous j rows.  calc_coeff gives a
// thread one point; the tendency kernels give it 8 (diffu_x) or 4 (diffu_d)
// rows of a tile whose values, with a halo of 2, are staged once in LDS.  calc_coeff forms xkd from the
// four surrounding raw coefficients, recomputing them rather than reading
// them back, so one launch writes xkc, xkd and xkcf with no ordering between
// threads.  The reference's quirks are kept: the del2 ring of idiffu = 1 is
// applied side by side (the corners twice), diffu_x3df runs its left / right
// ring over 1..kz and everything else over 1..kz+1, and idiffu = 3 diffuses
// the column j = jci2 (jdi2) only, with index clamps to the cross (dot)
// extent.  At level kz+1 of diffu_x3df under idiffu = 3 the reference reads
// xkc one level past its allocation; that level's coefficient is taken from
// xkcf, which under idiffu = 3 holds the same diff_6th_coef * psb as every
// level of xkc.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mlc_internal.h"

namespace mlc {
namespace diff {

constexpr int DTJ = 64, DTI = 4;
// the tendency kernels give each thread RPT rows of a DTJ x (DTI RPT) tile,
// so a block reads a (DTJ + 4) x (DTI RPT + 4) halo tile for DTJ x DTI RPT
// outputs: 8 rows for diffu_x (one tile), 4 for diffu_d (two tiles)
constexpr int RPT_X = 8, RPT_D = 4;
constexpr int RPT_C = 4;      // calc_coeff: rows per thread of its 64 x 16 tile
constexpr double kVonkar = 0.4;
// Main/mod_diffusion.F90:65-83
constexpr double o4_c1 = 4.0 / 6.0, o4_c2 = 1.0 / 6.0, o4_c3 = -20.0 / 6.0;
constexpr double z4_c1 = 1.0, z4_c2 = -4.0, z4_c3 = 12.0;
constexpr double h4_c1 = 10.0, h4_c2 = -5.0, h4_c3 = 1.0;
constexpr double kDiff6thFactor = 0.12;

struct Coeff {
  Dims d;
  const double *__restrict__ ud, *__restrict__ vd, *__restrict__ wx;
  const double *__restrict__ pc, *__restrict__ pd, *__restrict__ hgfact;
  double *__restrict__ xkc, *__restrict__ xkd, *__restrict__ xkcf;
  double xkhmax, dydc, coef6, rdxsq;
  int idiffu, idynamic;
};

// Smagorinsky coefficient before scaling (:195-226) at a cross point
__device__ __forceinline__ double raw_coeff(const Coeff &a, int j, int i, int k) {
  const Dims d = a.d;
  const long long c00 = d.at(j, i, k), c10 = c00 + 1, c01 = c00 + d.pj, c11 = c01 + 1;
  const double *ud = a.ud, *vd = a.vd;
  const double dudx = ud[c10] + ud[c11] - ud[c00] - ud[c01];
  const double dvdx = vd[c10] + vd[c11] - vd[c00] - vd[c01];
  const double dudy = ud[c01] + ud[c11] - ud[c00] - ud[c10];
  const double dvdy = vd[c01] + vd[c11] - vd[c00] - vd[c10];
  double duv;
  if (a.idynamic == 1) {
    duv = sqrt((dudx - dvdy) * (dudx - dvdy) + (dvdx + dudy) * (dvdx + dudy));
  } else {
    const double dwdz = a.wx[c00] - a.wx[c00 + d.plane];
    double s = (dudx - dvdy) * (dudx - dvdy) + (dvdx + dudy) * (dvdx + dudy) - dwdz * dwdz;
    if (!(s > 0.0)) s = 0.0;
    duv = sqrt(s);
  }
  const double x = a.hgfact[d.at2(j, i)] + a.dydc * duv;
  return x < a.xkhmax ? x : a.xkhmax;
}

__global__ void __launch_bounds__(DTJ *DTI) k_coeff(Coeff a) {
  constexpr int TR = DTI * RPT_C;
  const Dims d = a.d;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int j0 = 1 + blockIdx.x * DTJ, i0 = 1 + blockIdx.y * TR;
  const int j = j0 + tx, k = 1 + blockIdx.z;
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  if (a.idiffu == 3) {                                                // :170-180
    for (int r = 0; r < RPT_C; ++r) {
      const int i = i0 + ty + DTI * r;
      if (j > jx || i > iy) continue;
      const long long c = d.at(j, i, k), c2 = d.at2(j, i);
      const bool cin = j >= 2 && j <= jx - 2 && i >= 2 && i <= iy - 2;   // jci x ici
      const bool din = j >= 2 && j <= jx - 1 && i >= 2 && i <= iy - 1;   // jdi x idi
      if (cin) {
        const double v = a.coef6 * a.pc[c2];
        a.xkc[c] = v;
        a.xkcf[c] = v;
        if (k == kz) a.xkcf[c + d.plane] = v;
      }
      if (din) a.xkd[c] = a.coef6 * a.pd[c2];
    }
    return;
  }
  // raw coefficients of the tile and of its west column / south row, once each
  __shared__ double s_raw[TR + 1][DTJ + 1];
  double rv[RPT_C];
#pragma unroll
  for (int r = 0; r < RPT_C; ++r) {
    const int i = i0 + ty + DTI * r;
    const bool ce = j <= jx - 1 && i <= iy - 1;                      // jce x ice
    rv[r] = ce ? raw_coeff(a, j, i, k) : 0.0;
    s_raw[ty + DTI * r + 1][tx + 1] = rv[r];
  }
  const int t = ty * DTJ + tx;
  if (t < DTJ + 1 + TR) {
    const int hj = t < DTJ + 1 ? j0 - 1 + t : j0 - 1;
    const int hi = t < DTJ + 1 ? i0 - 1 : i0 + (t - DTJ - 1);
    const bool ok = hj >= 1 && hi >= 1 && hj <= jx - 1 && hi <= iy - 1;
    s_raw[hi - (i0 - 1)][hj - (j0 - 1)] = ok ? raw_coeff(a, hj, hi, k) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RPT_C; ++r) {
    const int i = i0 + ty + DTI * r, ry = ty + DTI * r;
    if (j > jx || i > iy) continue;
    const long long c = d.at(j, i, k), c2 = d.at2(j, i);
    const bool cin = j >= 2 && j <= jx - 2 && i >= 2 && i <= iy - 2;   // jci x ici
    const bool din = j >= 2 && j <= jx - 1 && i >= 2 && i <= iy - 1;   // jdi x idi
    const bool ce = j <= jx - 1 && i <= iy - 1;                        // jce x ice
    const double rr = rv[r];
    if (ce) {
      if (cin) {
        const double v = rr * a.rdxsq * a.pc[c2];
        a.xkc[c] = v;
        a.xkcf[c + d.plane] = v;                                      // xkcf(k+1) = xkc(k)
        if (k == 1) a.xkcf[c] = v;                                    // xkcf(1) = xkc(1)
      } else {
        a.xkc[c] = rr;
      }
    }
    if (din) {                                                        // :235-238, :245-247
      const double sum = s_raw[ry + 1][tx + 1] + s_raw[ry][tx] + s_raw[ry + 1][tx] +
                         s_raw[ry][tx + 1];
      a.xkd[c] = 0.25 * sum * a.rdxsq * a.pd[c2];
    }
  }
}

// hgfact (:131-146): xkhz on jce x ice, reduced by the steepest terrain
// gradient on jci x ici when diffu_hgtf = 1
__global__ void __launch_bounds__(DTJ *DTI) k_hgfact(Dims d, const double *__restrict__ ht,
                                                     double *__restrict__ hg, double xkhz,
                                                     double dx, int hgtf) {
  const int j = 1 + blockIdx.x * DTJ + threadIdx.x;
  const int i = 1 + blockIdx.y * DTI + threadIdx.y;
  if (j > d.jx - 1 || i > d.iy - 1) return;
  const long long c = d.at2(j, i);
  double v = xkhz;
  if (hgtf == 1 && j >= 2 && j <= d.jx - 2 && i >= 2 && i <= d.iy - 2) {
    const double h = ht[c];
    const double hg1 = fabs((h - ht[c - d.pj]) / dx);
    const double hg2 = fabs((h - ht[c + d.pj]) / dx);
    const double hg3 = fabs((h - ht[c - 1]) / dx);
    const double hg4 = fabs((h - ht[c + 1]) / dx);
    double m = hg1;
    if (hg2 > m) m = hg2;
    if (hg3 > m) m = hg3;
    if (hg4 > m) m = hg4;
    const double hgmax = m * (1.0 / kEgrav) * 1.0e3;
    v = xkhz / (1.0 + hgmax * hgmax);
  }
  hg[c] = v;
}

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// A field with an optional division by the dot-point map factor; Q(j, i)
// is the value at column (j, i) of the current level.
struct Plane {
  const double *f;      // level base
  const double *m;      // msfd (2-D) or null
  int pj;
  __device__ __forceinline__ double operator()(int j, int i) const {
    const long long c = (long long)(i - 1) * pj + (j - 1);
    return m ? f[c] / m[c] : f[c];
  }
  __device__ __forceinline__ double raw(int j, int i) const {
    return f[(long long)(i - 1) * pj + (j - 1)];
  }
  __device__ __forceinline__ double scaled(int j, int i) const {
    const long long c = (long long)(i - 1) * pj + (j - 1);
    return f[c] / m[c];
  }
};

template <class Q>
__device__ __forceinline__ double del4(const Q &q, int j, int i) {
  return z4_c1 * (q(j + 2, i) + q(j - 2, i) + q(j, i + 2) + q(j, i - 2)) +
         z4_c2 * (q(j + 1, i) + q(j - 1, i) + q(j, i + 1) + q(j, i - 1)) + z4_c3 * (q(j, i));
}
template <class Q>
__device__ __forceinline__ double del2(const Q &q, int j, int i) {
  return z4_c1 * (q(j + 1, i) + q(j - 1, i) + q(j, i + 1) + q(j, i - 1)) + z4_c2 * (q(j, i));
}
template <class Q>
__device__ __forceinline__ double nine(const Q &q, int j, int i) {
  return o4_c1 * (q(j + 1, i) + q(j - 1, i) + q(j, i + 1) + q(j, i - 1)) +
         o4_c2 * (q(j + 1, i + 1) + q(j - 1, i - 1) + q(j - 1, i + 1) + q(j + 1, i - 1)) +
         o4_c3 * (q(j, i));
}

// flux-limited 6th-order operand (:399-460, :581-630).  `flux` gives the
// values the flux differences use, `test` those of the limiter tests; the
// indices are clamped to [1, nj] x [1, ni].
template <class QF, class QT>
__device__ __forceinline__ double flux6(const QF &flux, const QT &test, int j, int i, int nj,
                                        int ni) {
  const int im1 = imax(i - 1, 1), im2 = imax(i - 2, 1), im3 = imax(i - 3, 1);
  const int ip1 = imin(i + 1, ni), ip2 = imin(i + 2, ni), ip3 = imin(i + 3, ni);
  const int jm1 = imax(j - 1, 1), jm2 = imax(j - 2, 1), jm3 = imax(j - 3, 1);
  const int jp1 = imin(j + 1, nj), jp2 = imin(j + 2, nj), jp3 = imin(j + 3, nj);
  double x0 = h4_c1 * (flux(j, i) - flux(jm1, i)) + h4_c2 * (flux(jp1, i) - flux(jm2, i)) +
              h4_c3 * (flux(jp2, i) - flux(jm3, i));
  if (x0 * (test(j, i) - test(jm1, i)) <= 0.0) x0 = 0.0;
  double x1 = h4_c1 * (flux(jp1, i) - flux(j, i)) + h4_c2 * (flux(jp2, i) - flux(jm1, i)) +
              h4_c3 * (flux(jp3, i) - flux(jm2, i));
  if (x1 * (test(jp1, i) - test(j, i)) <= 0.0) x1 = 0.0;
  double y0 = h4_c1 * (flux(j, i) - flux(j, im1)) + h4_c2 * (flux(j, ip1) - flux(j, im2)) +
              h4_c3 * (flux(j, ip2) - flux(j, im3));
  if (y0 * (test(j, i) - test(j, im1)) <= 0.0) y0 = 0.0;
  double y1 = h4_c1 * (flux(j, ip1) - flux(j, i)) + h4_c2 * (flux(j, ip2) - flux(j, im1)) +
              h4_c3 * (flux(j, ip3) - flux(j, im2));
  if (y1 * (test(j, ip1) - test(j, i)) <= 0.0) y1 = 0.0;
  return (x1 - x0) + (y1 - y0);
}

struct RawQ {
  Plane p;
  __device__ __forceinline__ double operator()(int j, int i) const { return p.raw(j, i); }
};
struct ScaledQ {
  Plane p;
  __device__ __forceinline__ double operator()(int j, int i) const { return p.scaled(j, i); }
};

// A (DTJ + 4) x (rows + 4) tile of one level with a halo of 2, in LDS:
// T(j, i) is the value at column (j, i) (1-based global indices).
constexpr int kTW = DTJ + 4;
struct Tile {
  const double (*s)[kTW];
  int jm, im;                                                         // j0 - 2, i0 - 2
  __device__ __forceinline__ double operator()(int j, int i) const { return s[i - im][j - jm]; }
};

// Fill a tile from a level base: f / msfd when m is set, f otherwise;
// points outside [1, nj] x [1, ni] hold zero (never read by a stencil).
template <int TH>
__device__ __forceinline__ void fill_tile(double (*s)[kTW], const double *f, const double *m,
                                          int pj, int j0, int i0, int nj, int ni) {
  const int t = threadIdx.y * DTJ + threadIdx.x;
  constexpr int kN = (kTW * TH + DTJ * DTI - 1) / (DTJ * DTI);
#pragma unroll
  for (int it = 0; it < kN; ++it) {
    const int e = t + it * DTJ * DTI;
    if (e >= kTW * TH) break;
    const int r = e / kTW, q = e - r * kTW;
    const int jj = j0 - 2 + q, ii = i0 - 2 + r;
    double v = 0.0;
    if (jj >= 1 && jj <= nj && ii >= 1 && ii <= ni) {
      const long long c = (long long)(ii - 1) * pj + (jj - 1);
      v = m ? f[c] / m[c] : f[c];
    }
    s[r][q] = v;
  }
}

// diffu_x: form 0 x3d (xkc), 1 x3df (fac * xkcf, kz+1 levels), 2 x4d3d
// (fac * xkc); species n1..n2 of F / FTEN (block stride kz+1 levels)
struct XArgs {
  Dims d;
  const double *__restrict__ f;
  double *__restrict__ ften;
  const double *__restrict__ xkc, *__restrict__ xkcf, *__restrict__ msfd;
  double fac;
  int form, idiffu, nk, n1;
};

__global__ void __launch_bounds__(DTJ *DTI) k_diffu_x(XArgs a) {
  const Dims d = a.d;
  constexpr int TR = DTI * RPT_X;
  const int j0 = 1 + blockIdx.x * DTJ, i0 = 1 + blockIdx.y * TR;
  const int j = j0 + threadIdx.x;
  const int k = 1 + (int)(blockIdx.z % (unsigned)a.nk);
  const int n = a.n1 + (int)(blockIdx.z / (unsigned)a.nk);
  const int jx = d.jx, iy = d.iy, kz = d.kz;
  const long long blk = (long long)(n - 1) * (kz + 1) * d.plane;
  const long long lev = (long long)(k - 1) * d.plane;
  const Plane q{a.f + blk + lev, a.msfd, d.pj};
  __shared__ double sf[TR + 4][kTW];
  if (a.idiffu != 3) {                                               // the field's tile
    fill_tile<TR + 4>(sf, a.f + blk + lev, nullptr, d.pj, j0, i0, jx - 1, iy - 1);
    __syncthreads();
  }
  const Tile tf{sf, j0 - 2, i0 - 2};
  for (int r = 0; r < RPT_X; ++r) {
    const int i = i0 + (int)threadIdx.y + DTI * r;
    const bool act = j >= 2 && j <= jx - 2 && i >= 2 && i <= iy - 2;  // jci x ici
    if (!act) continue;
    if (a.idiffu == 3 && j != jx - 2) continue;                      // column jci2 only
    const long long c = blk + lev + d.at2(j, i);
    const long long kc = lev + d.at2(j, i);                           // coefficient index
    double v = a.ften[c];
    if (a.idiffu == 1) {
      const double kf = a.form == 0 ? a.xkc[kc] : a.fac * (a.form == 1 ? a.xkcf : a.xkc)[kc];
      if (j >= 3 && j <= jx - 3 && i >= 3 && i <= iy - 3) {
        v = v - kf * del4(tf, j, i);
      } else {
        const double t = kf * del2(tf, j, i);
        if (k <= kz && j == 2) v = v + t;                             // left
        if (k <= kz && j == jx - 2) v = v + t;                        // right
        if (i == 2) v = v + t;                                        // bottom
        if (i == iy - 2) v = v + t;                                   // top
      }
    } else if (a.idiffu == 2) {
      const double kf = a.form == 0 ? a.xkc[kc] : a.fac * (a.form == 1 ? a.xkcf : a.xkc)[kc];
      v = v + kf * nine(tf, j, i);
    } else {
      const double xkv = (a.form == 1 && k == kz + 1) ? a.xkcf[kc] : a.xkc[kc];
      const double kf = a.form == 1 ? a.fac * xkv : xkv;
      v = v + kf * flux6(RawQ{q}, ScaledQ{q}, j, i, jx - 1, iy - 1);
    }
    a.ften[c] = v;
  }
}

// diffu_d: uten / vten from u / msfd and v / msfd at dot points
struct DArgs {
  Dims d;
  const double *__restrict__ u, *__restrict__ v, *__restrict__ xkd, *__restrict__ msfd;
  double *__restrict__ uten, *__restrict__ vten;
  int idiffu;
};

template <class Q>
__device__ __forceinline__ double dot_term(int idiffu, const Q &q, int j, int i, int jx, int iy,
                                           double kd, double t) {
  if (idiffu == 1) {
    if (j >= 3 && j <= jx - 2 && i >= 3 && i <= iy - 2) return t - kd * del4(q, j, i);
    const double s = kd * del2(q, j, i);
    if (j == 2) t = t + s;
    if (j == jx - 1) t = t + s;
    if (i == 2) t = t + s;
    if (i == iy - 1) t = t + s;
    return t;
  }
  return t + kd * nine(q, j, i);
}

__global__ void __launch_bounds__(DTJ *DTI) k_diffu_d(DArgs a) {
  const Dims d = a.d;
  constexpr int TR = DTI * RPT_D;
  const int j0 = 1 + blockIdx.x * DTJ, i0 = 1 + blockIdx.y * TR;
  const int j = j0 + threadIdx.x, k = 1 + blockIdx.z;
  const int jx = d.jx, iy = d.iy;
  const long long lev = (long long)(k - 1) * d.plane;
  if (a.idiffu == 3) {                                               // column jdi2 only
    if (j != jx - 1) return;
    const Plane pu{a.u + lev, a.msfd, d.pj}, pv{a.v + lev, a.msfd, d.pj};
    for (int r = 0; r < RPT_D; ++r) {
      const int i = i0 + (int)threadIdx.y + DTI * r;
      if (i < 2 || i > iy - 1) continue;                              // idi
      const long long c = lev + d.at2(j, i);
      const double kd = a.xkd[c];
      a.uten[c] = a.uten[c] + kd * flux6(ScaledQ{pu}, ScaledQ{pu}, j, i, jx, iy);
      a.vten[c] = a.vten[c] + kd * flux6(ScaledQ{pv}, ScaledQ{pv}, j, i, jx, iy);
    }
    return;
  }
  // u / msfd and v / msfd of the tile and its halo, each divided once
  __shared__ double su[TR + 4][kTW], sv[TR + 4][kTW];
  fill_tile<TR + 4>(su, a.u + lev, a.msfd, d.pj, j0, i0, jx, iy);
  fill_tile<TR + 4>(sv, a.v + lev, a.msfd, d.pj, j0, i0, jx, iy);
  __syncthreads();
  for (int r = 0; r < RPT_D; ++r) {
    const int i = i0 + (int)threadIdx.y + DTI * r;
    const bool act = j >= 2 && j <= jx - 1 && i >= 2 && i <= iy - 1;  // jdi x idi
    if (!act) continue;
    const long long c = lev + d.at2(j, i);
    const double kd = a.xkd[c];
    a.uten[c] = dot_term(a.idiffu, Tile{su, j0 - 2, i0 - 2}, j, i, jx, iy, kd, a.uten[c]);
    a.vten[c] = dot_term(a.idiffu, Tile{sv, j0 - 2, i0 - 2}, j, i, jx, iy, kd, a.vten[c]);
  }
}

}  // namespace diff
}  // namespace mlc

struct mlc_diff {
  mlc_diff_desc desc{};
  mlc::Dims d{};
  hipStream_t stream = nullptr;
  double *ht = nullptr, *msfd = nullptr, *pc = nullptr, *pd = nullptr, *ud = nullptr,
         *vd = nullptr, *wx = nullptr, *uten = nullptr, *vten = nullptr, *f = nullptr,
         *ften = nullptr, *hgfact = nullptr, *xkc = nullptr, *xkd = nullptr, *xkcf = nullptr;
  double xkhz = 0, xkhmax = 0, dydc = 0, coef6 = 0;
  bool have_ht = false, inited = false, have_coeff = false;
};

namespace {

using mlc::fail;

struct DField {
  double **p;
  int nk, nspec;
};

bool dfield(mlc_diff *h, int id, DField *o) {
  const int kz = h->d.kz, ns = h->desc.nspec;
  switch (id) {
    case MLC_DF_HT: *o = {&h->ht, 1, 1}; return true;
    case MLC_DF_MSFD: *o = {&h->msfd, 1, 1}; return true;
    case MLC_DF_PC: *o = {&h->pc, 1, 1}; return true;
    case MLC_DF_PD: *o = {&h->pd, 1, 1}; return true;
    case MLC_DF_UD: *o = {&h->ud, kz, 1}; return true;
    case MLC_DF_VD: *o = {&h->vd, kz, 1}; return true;
    case MLC_DF_WX: *o = {&h->wx, kz + 1, 1}; return true;
    case MLC_DF_UTEN: *o = {&h->uten, kz, 1}; return true;
    case MLC_DF_VTEN: *o = {&h->vten, kz, 1}; return true;
    case MLC_DF_F: *o = {&h->f, kz + 1, ns}; return true;
    case MLC_DF_FTEN: *o = {&h->ften, kz + 1, ns}; return true;
    case MLC_DF_HGFACT: *o = {&h->hgfact, 1, 1}; return true;
    case MLC_DF_XKC: *o = {&h->xkc, kz, 1}; return true;
    case MLC_DF_XKD: *o = {&h->xkd, kz, 1}; return true;
    case MLC_DF_XKCF: *o = {&h->xkcf, kz + 1, 1}; return true;
    default: return false;
  }
}

void free_diff(mlc_diff *h) {
  double **ps[] = {&h->ht, &h->msfd, &h->pc, &h->pd, &h->ud, &h->vd, &h->wx, &h->uten,
                   &h->vten, &h->f, &h->ften, &h->hgfact, &h->xkc, &h->xkd, &h->xkcf};
  for (auto p : ps)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

int copy(mlc_diff *h, int id, double *host, const int lb[4], const int ub[4], bool to_dev) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  if (!host || !lb || !ub) return fail(MLC_E_ARG, "null host pointer or bounds");
  DField fd;
  if (!dfield(h, id, &fd)) return fail(MLC_E_ARG, "unknown diffusion field id");
  const int ext[4] = {h->d.jx, h->d.iy, fd.nk, fd.nspec};
  for (int m = 0; m < 4; ++m)
    if (lb[m] < 1 || ub[m] < lb[m] || ub[m] > ext[m]) {
      char buf[160];
      snprintf(buf, sizeof buf, "diffusion field %d: box dim %d [%d:%d] outside [1:%d]", id,
               m + 1, lb[m], ub[m], ext[m]);
      return fail(MLC_E_BOUNDS, buf);
    }
  const long long nj = ub[0] - lb[0] + 1, ni = ub[1] - lb[1] + 1, nk = ub[2] - lb[2] + 1,
                  nn = ub[3] - lb[3] + 1;
  const mlc::Dims d = h->d;
  for (long long n = 0; n < nn; ++n)
    for (long long k = 0; k < nk; ++k) {
      double *dev = *fd.p + (long long)(lb[3] - 1 + n) * fd.nk * d.plane +
                    d.at(lb[0], lb[1], (int)(lb[2] + k));
      double *hp = host + ((n * nk + k) * ni) * nj;
      if (to_dev)
        MLC_HIP(hipMemcpy2DAsync(dev, d.pj * sizeof(double), hp, nj * sizeof(double),
                                 nj * sizeof(double), ni, hipMemcpyHostToDevice, h->stream));
      else
        MLC_HIP(hipMemcpy2DAsync(hp, nj * sizeof(double), dev, d.pj * sizeof(double),
                                 nj * sizeof(double), ni, hipMemcpyDeviceToHost, h->stream));
    }
  MLC_HIP(hipStreamSynchronize(h->stream));
  if (to_dev && id == MLC_DF_HT) h->have_ht = true;
  return MLC_OK;
}

dim3 plane_grid(const mlc::Dims &d, unsigned nz) {
  return dim3((unsigned)((d.jx + mlc::diff::DTJ - 1) / mlc::diff::DTJ),
              (unsigned)((d.iy + mlc::diff::DTI - 1) / mlc::diff::DTI), nz);
}
// the tendency kernels: DTJ x (DTI rpt) tiles
dim3 tile_grid(const mlc::Dims &d, unsigned nz, int rpt) {
  const int rows = mlc::diff::DTI * rpt;
  return dim3((unsigned)((d.jx + mlc::diff::DTJ - 1) / mlc::diff::DTJ),
              (unsigned)((d.iy + rows - 1) / rows), nz);
}

}  // namespace

extern "C" {

int mlc_diff_create(const mlc_diff_desc *desc, mlc_diff **out) {
  if (!desc || !out) return fail(MLC_E_ARG, "null descriptor or output");
  *out = nullptr;
  const mlc_diff_desc &g = *desc;
  if (g.jx < 8 || g.iy < 8 || g.kz < 1 || g.idiffu < 1 || g.idiffu > 3 || g.idynamic < 1 ||
      g.idynamic > 2 || g.diffu_hgtf < 0 || g.diffu_hgtf > 1 || g.nspec < 1 ||
      g.reserved != 0 || !(g.dx > 0.0) || !(g.dtsec > 0.0))
    return fail(MLC_E_ARG, "bad diffusion descriptor");
  mlc_diff *h = new (std::nothrow) mlc_diff;
  if (!h) return fail(MLC_E_ARG, "out of host memory");
  h->desc = g;
  h->d = mlc::make_dims(g.jx, g.iy, g.kz, 0);
  auto bail = [&](int rc) {
    free_diff(h);
    delete h;
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(mlc::hip_check(e, "hipStreamCreateWithFlags"));
  const long long P = h->d.plane, kz = g.kz;
  struct { double **p; long long n; } al[] = {
      {&h->ht, P}, {&h->msfd, P}, {&h->pc, P}, {&h->pd, P}, {&h->ud, P * kz},
      {&h->vd, P * kz}, {&h->wx, P * (kz + 1)}, {&h->uten, P * kz}, {&h->vten, P * kz},
      {&h->f, P * (kz + 1) * g.nspec}, {&h->ften, P * (kz + 1) * g.nspec}, {&h->hgfact, P},
      {&h->xkc, P * kz}, {&h->xkd, P * kz}, {&h->xkcf, P * (kz + 1)}};
  for (auto &a : al) {
    if ((e = hipMalloc(a.p, (size_t)a.n * sizeof(double))) != hipSuccess)
      return bail(mlc::hip_check(e, "hipMalloc"));
    if ((e = hipMemset(*a.p, 0, (size_t)a.n * sizeof(double))) != hipSuccess)
      return bail(mlc::hip_check(e, "hipMemset"));
  }
  // initialize_diffusion's scalars (:110-125, :152-153)
  const double dxsq = g.dx * g.dx;
  h->xkhmax = dxsq / (64.0 * g.dtsec);
  h->dydc = g.adyndif * mlc::diff::kVonkar * mlc::diff::kVonkar * g.dx * 0.25;
  if (g.idynamic == 1) {
    h->xkhz = g.ckh * 1.5e-3 * dxsq / g.dtsec;
  } else {
    h->xkhz = g.ckh * g.dx;
    h->xkhmax = 2.0 * h->xkhmax;
  }
  h->coef6 = mlc::diff::kDiff6thFactor * 0.015625 / (2.0 * g.dtsec);
  *out = h;
  return MLC_OK;
}

int mlc_diff_destroy(mlc_diff *h) {
  if (!h) return MLC_OK;
  free_diff(h);
  delete h;
  return MLC_OK;
}

int mlc_diff_put(mlc_diff *h, int field, const double *host, const int lb[4],
                 const int ub[4]) {
  return copy(h, field, const_cast<double *>(host), lb, ub, true);
}

int mlc_diff_get(mlc_diff *h, int field, double *host, const int lb[4], const int ub[4]) {
  return copy(h, field, host, lb, ub, false);
}

int mlc_diff_initialize(mlc_diff *h) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  if (h->desc.idiffu < 3) {
    if (h->desc.diffu_hgtf == 1 && !h->have_ht)
      return fail(MLC_E_STATE, "MLC_DF_HT has not been set (diffu_hgtf = 1)");
    hipLaunchKernelGGL(mlc::diff::k_hgfact, plane_grid(h->d, 1),
                       dim3(mlc::diff::DTJ, mlc::diff::DTI), 0, h->stream, h->d, h->ht,
                       h->hgfact, h->xkhz, h->desc.dx, h->desc.diffu_hgtf);
    MLC_HIP(hipGetLastError());
  }
  h->inited = true;
  return MLC_OK;
}

int mlc_diff_calc_coeff(mlc_diff *h) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_diff_initialize has not been called");
  mlc::diff::Coeff a{h->d, h->ud, h->vd, h->wx, h->pc, h->pd, h->hgfact, h->xkc, h->xkd,
                     h->xkcf, h->xkhmax, h->dydc, h->coef6,
                     1.0 / (h->desc.dx * h->desc.dx), h->desc.idiffu, h->desc.idynamic};
  hipLaunchKernelGGL(mlc::diff::k_coeff, tile_grid(h->d, (unsigned)h->d.kz, mlc::diff::RPT_C),
                     dim3(mlc::diff::DTJ, mlc::diff::DTI), 0, h->stream, a);
  MLC_HIP(hipGetLastError());
  h->have_coeff = true;
  return MLC_OK;
}

int mlc_diff_d(mlc_diff *h) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  if (!h->have_coeff) return fail(MLC_E_STATE, "mlc_diff_calc_coeff has not been called");
  mlc::diff::DArgs a{h->d, h->ud, h->vd, h->xkd, h->msfd, h->uten, h->vten, h->desc.idiffu};
  hipLaunchKernelGGL(mlc::diff::k_diffu_d, tile_grid(h->d, (unsigned)h->d.kz, mlc::diff::RPT_D),
                     dim3(mlc::diff::DTJ, mlc::diff::DTI), 0, h->stream, a);
  MLC_HIP(hipGetLastError());
  return MLC_OK;
}

int mlc_diff_x(mlc_diff *h, int form, int n1, int n2, double fac) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  if (!h->have_coeff) return fail(MLC_E_STATE, "mlc_diff_calc_coeff has not been called");
  if (form < MLC_DIFFU_X3D || form > MLC_DIFFU_X4D) return fail(MLC_E_ARG, "bad form");
  if (n1 < 1 || n2 < n1 || n2 > h->desc.nspec) return fail(MLC_E_ARG, "bad species range");
  const int nk = form == MLC_DIFFU_X3DF ? h->d.kz + 1 : h->d.kz;
  if ((long long)nk * (n2 - n1 + 1) > 65535)
    return fail(MLC_E_ARG, "levels x species exceed one launch (65535)");
  mlc::diff::XArgs a{h->d, h->f, h->ften, h->xkc, h->xkcf, h->msfd, fac, form,
                     h->desc.idiffu, nk, n1};
  hipLaunchKernelGGL(mlc::diff::k_diffu_x,
                     tile_grid(h->d, (unsigned)(nk * (n2 - n1 + 1)), mlc::diff::RPT_X),
                     dim3(mlc::diff::DTJ, mlc::diff::DTI), 0, h->stream, a);
  MLC_HIP(hipGetLastError());
  return MLC_OK;
}

int mlc_diff_synchronize(mlc_diff *h) {
  if (!h) return fail(MLC_E_ARG, "null diffusion handle");
  MLC_HIP(hipStreamSynchronize(h->stream));
  return MLC_OK;
}

}  // extern "C"
