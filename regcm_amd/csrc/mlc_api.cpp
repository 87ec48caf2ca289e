// C-ABI of libmoloch_hip: handle lifecycle, host<->device transfers of
// Fortran-bounded boxes, init_moloch's derived metrics and the step
// orchestration of moloch() (Main/mod_moloch.F90:312-446).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "mlc_internal.h"

namespace mlc {

static thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  return fail(MLC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static hipEvent_t prof_take(mlc_handle *h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

void prof_events(mlc_handle *h, int slot, hipEvent_t *a, hipEvent_t *b) {
  *a = *b = nullptr;
  if (!h->prof || !(h->prof_mask & (1u << slot))) return;
  hipEvent_t x = prof_take(h);
  if (x == nullptr) return;
  hipEvent_t y = prof_take(h);
  if (y == nullptr) {
    h->pool.push_back(x);
    return;
  }
  *a = x;
  *b = y;
}

void prof_push(mlc_handle *h, int slot, hipEvent_t a, hipEvent_t b) {
  if (a != nullptr) h->pending.push_back({slot, a, b});
}

// Resolve pending event pairs into per-slot totals.
static int prof_resolve(mlc_handle *h) {
  if (h->pending.empty()) return MLC_OK;
  MLC_HIP(hipStreamSynchronize(h->stream));
  for (auto &e : h->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      h->launches[e.slot] += 1;
      h->ms[e.slot] += ms;
    }
    h->pool.push_back(e.a);
    h->pool.push_back(e.b);
  }
  h->pending.clear();
  return MLC_OK;
}

// gzita: Share/mod_zita.F90:109-119;  model_zitaf/h: :54-84
static double gzita(double zita, double ztop, double a0) {
  const double ratio = zita / ztop;
  return (((0.0 - 1.0 * a0) * ratio - (3.0 - 2.0 * a0) * (ratio * ratio)) +
          (2.0 - 1.0 * a0) * (ratio * ratio * ratio)) + 1.0;
}

// gzitap (Share/mod_zita.F90:122-132): ratio**0 is 1, ratio**2 is ratio*ratio
static double gzitap(double zita, double ztop, double a0) {
  const double ratio = zita / ztop;
  return (((0.0 - 1.0 * a0) * 1.0 - (6.0 - 4.0 * a0) * ratio) +
          (6.0 - 3.0 * a0) * (ratio * ratio)) / ztop;
}

// bzitap = zfz(ztop, zh) * exp(zita/zh) / zh, zfz = ztop/(exp(ztop/zh) - 1)
// (:88-93, :102-107), with the C library's exp as the Fortran runtime's
static double bzitap(double zita, double ztop, double zh) {
  const double zfz = ztop / (std::exp(ztop / zh) - 1.0);
  return zfz * std::exp(zita / zh) / zh;
}

struct FieldInfo {
  Buf *buf = nullptr;        // bulk-type device array (null for special fields)
  double **dptr = nullptr;   // always-double device array
  int adv_slot = -1;         // advected field slot or -1
  int nk = 0;                // levels
  int nspec = 1;             // species (qx) or 1
  bool stat = false;         // static metric
};

static bool field_info(mlc_handle *h, int f, FieldInfo *fi) {
  const int kz = h->d.kz;
  *fi = FieldInfo{};
  fi->nk = kz;
  switch (f) {
    case MLC_U: fi->buf = &h->u; break;
    case MLC_V: fi->buf = &h->v; break;
    case MLC_W: fi->dptr = &h->w; fi->nk = kz + 1; break;
    case MLC_PAI: fi->adv_slot = A_PAI; break;
    case MLC_TETAV: fi->adv_slot = A_TETAV; break;
    case MLC_T: fi->buf = &h->t; break;
    case MLC_TVIRT: fi->buf = &h->tvirt; break;
    case MLC_QX: fi->adv_slot = A_QX0; fi->nspec = h->g.nqx; break;
    case MLC_UX: fi->adv_slot = A_UX; break;
    case MLC_VX: fi->adv_slot = A_VX; break;
    case MLC_P: fi->buf = &h->p; break;
    case MLC_RHO: fi->buf = &h->rho; break;
    case MLC_QSAT: fi->buf = &h->qsat; break;
    case MLC_PS: fi->buf = &h->ps; fi->nk = 1; break;
    case MLC_TTEN: fi->buf = &h->tten; break;
    case MLC_UTEN: fi->buf = &h->uten; break;
    case MLC_VTEN: fi->buf = &h->vten; break;
    case MLC_QXTEN: fi->buf = &h->qxten; fi->nspec = h->g.nqx; break;
    case MLC_TKE:
      if (h->s_tkex < 0) return false;
      fi->buf = &h->tke; fi->nk = kz + 1; break;
    case MLC_TKETEN:
      if (h->s_tkex < 0) return false;
      fi->buf = &h->tketen; fi->nk = kz + 1; break;
    case MLC_TRAC:
      if (h->s_tr0 < 0) return false;
      fi->adv_slot = h->s_tr0; fi->nspec = h->g.ntr; break;
    case MLC_CHITEN:
      if (h->s_tr0 < 0) return false;
      fi->buf = &h->chiten; fi->nspec = h->g.ntr; break;
    case MLC_FMZ: fi->buf = &h->fmz; fi->stat = true; break;
    case MLC_FMZF: fi->buf = &h->fmzf; fi->nk = kz + 1; fi->stat = true; break;
    case MLC_RFMZU: fi->buf = &h->rfmzu; fi->stat = true; break;
    case MLC_RFMZV: fi->buf = &h->rfmzv; fi->stat = true; break;
    case MLC_BDYWTU: fi->buf = &h->bdywtu; fi->stat = true; break;
    case MLC_BDYWTV: fi->buf = &h->bdywtv; fi->stat = true; break;
    case MLC_BDYWTW: fi->buf = &h->bdywtw; fi->stat = true; break;
    case MLC_ZETA: fi->buf = &h->zeta; fi->stat = true; break;
    case MLC_DZ: fi->buf = &h->dz; fi->stat = true; break;
    case MLC_MX: fi->buf = &h->mx; fi->nk = 1; fi->stat = true; break;
    case MLC_MU: fi->buf = &h->mu; fi->nk = 1; fi->stat = true; break;
    case MLC_MV: fi->buf = &h->mv; fi->nk = 1; fi->stat = true; break;
    case MLC_HX: fi->buf = &h->hx; fi->nk = 1; fi->stat = true; break;
    case MLC_HY: fi->buf = &h->hy; fi->nk = 1; fi->stat = true; break;
    case MLC_CORU: fi->buf = &h->coru; fi->nk = 1; fi->stat = true; break;
    case MLC_CORV: fi->buf = &h->corv; fi->nk = 1; fi->stat = true; break;
    case MLC_RLAT: fi->nk = 1; fi->stat = true; break;
    case MLC_FFILT: fi->nk = kz; fi->stat = true; break;
    case MLC_HT: fi->dptr = &h->geo[0]; fi->nk = 1; fi->stat = true; break;
    case MLC_HTU: fi->dptr = &h->geo[1]; fi->nk = 1; fi->stat = true; break;
    case MLC_HTV: fi->dptr = &h->geo[2]; fi->nk = 1; fi->stat = true; break;
    case MLC_ZETAF: fi->dptr = &h->sl.zetaf; fi->nk = kz + 1; fi->stat = true; break;
    case MLC_XLAT: fi->dptr = &h->sl.xlat; fi->nk = 1; fi->stat = true; break;
    case MLC_SL_PF: fi->dptr = &h->sl.pf; fi->nk = kz + 1; break;
    case MLC_SL_TH: fi->dptr = &h->sl.th; break;
    case MLC_SL_RH: fi->dptr = &h->sl.rh; break;
    case MLC_SL_OMEGA: fi->dptr = &h->sl.omega; break;
    case MLC_SL_RHOX: fi->dptr = &h->sl.rhox; fi->nk = 1; break;
    case MLC_SL_TP: fi->dptr = &h->sl.tp; fi->nk = 1; break;
    case MLC_SL_TH700: fi->dptr = &h->sl.th700; fi->nk = 1; break;
    case MLC_SL_PTROP: fi->dptr = &h->sl.ptrop; fi->nk = 1; break;
    case MLC_SL_KTROP: fi->dptr = &h->sl.ktrop; fi->nk = 1; break;
    case MLC_SL_KMXPBL: fi->dptr = &h->sl.kmxpbl; fi->nk = 1; break;
    default: {
      // lateral-boundary inputs (after mlc_lbc_configure)
      if (!h->lbc.on) return false;
      double **bp[] = {&h->lbc.ps0, &h->lbc.ps1, &h->lbc.u0, &h->lbc.u1, &h->lbc.v0,
                       &h->lbc.v1, &h->lbc.t0, &h->lbc.t1, &h->lbc.pai0, &h->lbc.pai1,
                       &h->lbc.qv0, &h->lbc.qv1, &h->lbc.qc0, &h->lbc.qc1, &h->lbc.qi0,
                       &h->lbc.qi1, &h->lbc.xfu, &h->lbc.xfv, &h->lbc.xfc};
      if (f >= MLC_BND_PS0 && f <= MLC_RELAX_C) {
        fi->dptr = bp[f - MLC_BND_PS0];
        if (f == MLC_BND_PS0 || f == MLC_BND_PS1) fi->nk = 1;
        return true;
      }
      if (f == MLC_TNUDGE) {
        fi->nk = kz;
        return true;
      }
      return false;
    }
  }
  return true;
}

static int check_box(const mlc_handle *h, const FieldInfo &fi, int f,
                     const int lb[4], const int ub[4]) {
  if (!lb || !ub) return fail(MLC_E_ARG, "null bounds");
  int ext[4] = {h->d.jx, h->d.iy, fi.nk, fi.nspec};
  if (f == MLC_RLAT) { ext[0] = 1; ext[1] = h->d.iy; }
  if (f == MLC_FFILT || f == MLC_TNUDGE) { ext[0] = 1; ext[1] = 1; }
  for (int m = 0; m < 4; ++m)
    if (lb[m] < 1 || ub[m] < lb[m] || ub[m] > ext[m]) {
      char buf[160];
      snprintf(buf, sizeof buf,
               "field %d: box dim %d [%d:%d] outside extent [1:%d]", f, m + 1,
               lb[m], ub[m], ext[m]);
      return fail(MLC_E_BOUNDS, buf);
    }
  return MLC_OK;
}

// Device base (as bytes) and element size of species n (0-based) of a field.
static char *field_base(mlc_handle *h, const FieldInfo &fi, int n, int *es) {
  if (fi.adv_slot >= 0) {
    const Buf &b = h->cur_adv(fi.adv_slot + n);
    *es = b.es;
    return static_cast<char *>(b.p);
  }
  if (fi.dptr) {
    *es = 8;
    return reinterpret_cast<char *>(*fi.dptr + (long long)n * h->d.plane * fi.nk);
  }
  *es = fi.buf->es;
  return static_cast<char *>(fi.buf->p) +
         (long long)n * h->d.plane * fi.nk * fi.buf->es;
}

// Pinned host staging of at least `bytes` (grown on demand, kept).
static int stage_reserve(mlc_handle *h, size_t bytes) {
  if (h->pinned_len >= bytes) return MLC_OK;
  if (h->pinned) {
    (void)hipHostFree(h->pinned);
    h->pinned = nullptr;
    h->pinned_len = 0;
  }
  MLC_HIP(hipHostMalloc(&h->pinned, bytes, hipHostMallocDefault));
  h->pinned_len = bytes;
  return MLC_OK;
}

// The mkslice inputs and outputs, allocated on first use: zero-filled, with
// ptrop at the 100 hPa that init sets for irceideal = 1
// (Main/mod_init.F90:968-972) and ktrop / kmxpbl at kz.
static int slice_alloc(mlc_handle *h) {
  if (h->sl.on) return MLC_OK;
  const long long P = h->d.plane, kz = h->d.kz;
  struct { double **p; long long n; double v; } a[] = {
      {&h->sl.zetaf, P * (kz + 1), 0.0}, {&h->sl.xlat, P, 0.0},
      {&h->sl.pf, P * (kz + 1), 0.0},    {&h->sl.th, P * kz, 0.0},
      {&h->sl.rh, P * kz, 0.0},          {&h->sl.omega, P * kz, 0.0},
      {&h->sl.rhox, P, 0.0},             {&h->sl.tp, P, 0.0},
      {&h->sl.th700, P, 0.0},            {&h->sl.ptrop, P, 10000.0},
      {&h->sl.ktrop, P, (double)kz},     {&h->sl.kmxpbl, P, (double)kz}};
  for (auto &e : a) {
    MLC_HIP(hipMalloc(e.p, (size_t)e.n * sizeof(double)));
    std::vector<double> init((size_t)e.n, e.v);
    MLC_HIP(hipMemcpy(*e.p, init.data(), (size_t)e.n * sizeof(double), hipMemcpyHostToDevice));
  }
  h->sl.on = true;
  return MLC_OK;
}

static bool is_slice_field(int f) {
  return f == MLC_ZETAF || f == MLC_XLAT || (f >= MLC_SL_PF && f <= MLC_SL_KMXPBL);
}

static int copy_box(mlc_handle *h, int f, double *host, const int lb[4],
                    const int ub[4], bool to_dev) {
  FieldInfo fi;
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!host) return fail(MLC_E_ARG, "null host pointer");
  if (!field_info(h, f, &fi)) return fail(MLC_E_ARG, "unknown field id");
  int rc = check_box(h, fi, f, lb, ub);
  if (rc) return rc;
  if (is_slice_field(f)) {
    if ((rc = slice_alloc(h))) return rc;
    if (to_dev && f == MLC_ZETAF) h->sl.have_zetaf = true;
    if (to_dev && f == MLC_XLAT) h->sl.have_xlat = true;
  }
  const long long nj = ub[0] - lb[0] + 1, ni = ub[1] - lb[1] + 1,
                  nk = ub[2] - lb[2] + 1, nn = ub[3] - lb[3] + 1;
  if (f == MLC_TNUDGE) {                    // 1-D, device-resident
    const size_t off = (size_t)(lb[2] - 1), n = (size_t)nk;
    if (to_dev)
      MLC_HIP(hipMemcpy(h->lbc.tnudge + off, host, n * sizeof(double), hipMemcpyHostToDevice));
    else
      MLC_HIP(hipMemcpy(host, h->lbc.tnudge + off, n * sizeof(double), hipMemcpyDeviceToHost));
    return MLC_OK;
  }
  if (f == MLC_RLAT || f == MLC_FFILT) {   // host-resident 1-D inputs
    std::vector<double> &dst = (f == MLC_RLAT) ? h->h_rlat : h->h_ffilt;
    const int off = (f == MLC_RLAT) ? lb[1] - 1 : lb[2] - 1;
    const long long n = (f == MLC_RLAT) ? ni : nk;
    for (long long m = 0; m < n; ++m) {
      if (to_dev) dst[off + m] = host[m];
      else host[m] = dst[off + m];
    }
    if (to_dev) {
      h->have_static[f - MLC_FMZ] = true;
      h->inited = false;
    }
    return MLC_OK;
  }
  const long long nbox = nj * ni * nk * nn;
  const bool whole_planes = lb[0] == 1 && ub[0] == h->d.jx && lb[1] == 1 && ub[1] == h->d.iy;
  if (whole_planes) {
    // whole (j, i) planes: the device range of k = lb..ub is contiguous;
    // pack into pinned staging with the device pitch (pad columns zero)
    // and move each species with one copy
    const long long P = h->d.plane, pj = h->d.pj;
    for (long long n = 0; n < nn; ++n) {
      int es = 8;
      char *base = field_base(h, fi, (int)(lb[3] - 1 + n), &es);
      char *dev = base + h->d.at(1, 1, lb[2]) * es;
      const size_t bytes = (size_t)(nk * P * es);
      if ((rc = stage_reserve(h, bytes))) return rc;
      const double *hsrc = host + n * nk * ni * nj;
      if (to_dev) {
        if (es == 8) {
          double *st = static_cast<double *>(h->pinned);
          for (long long k = 0; k < nk; ++k)
            for (long long i = 0; i < ni; ++i) {
              double *row = st + k * P + i * pj;
              std::memcpy(row, hsrc + (k * ni + i) * nj, (size_t)nj * sizeof(double));
              for (long long j = nj; j < pj; ++j) row[j] = 0.0;
            }
        } else {
          float *st = static_cast<float *>(h->pinned);
          for (long long k = 0; k < nk; ++k)
            for (long long i = 0; i < ni; ++i) {
              float *row = st + k * P + i * pj;
              const double *src = hsrc + (k * ni + i) * nj;
              for (long long j = 0; j < nj; ++j) row[j] = (float)src[j];
              for (long long j = nj; j < pj; ++j) row[j] = 0.0f;
            }
        }
        MLC_HIP(hipMemcpyAsync(dev, h->pinned, bytes, hipMemcpyHostToDevice, h->stream));
        MLC_HIP(hipStreamSynchronize(h->stream));
      } else {
        MLC_HIP(hipMemcpyAsync(h->pinned, dev, bytes, hipMemcpyDeviceToHost, h->stream));
        MLC_HIP(hipStreamSynchronize(h->stream));
        double *hdst = host + n * nk * ni * nj;
        for (long long k = 0; k < nk; ++k)
          for (long long i = 0; i < ni; ++i) {
            double *dst = hdst + (k * ni + i) * nj;
            if (es == 8) {
              std::memcpy(dst, static_cast<double *>(h->pinned) + k * P + i * pj,
                          (size_t)nj * sizeof(double));
            } else {
              const float *row = static_cast<float *>(h->pinned) + k * P + i * pj;
              for (long long j = 0; j < nj; ++j) dst[j] = (double)row[j];
            }
          }
      }
    }
  } else {
    std::vector<float> stage;      // float staging for 4-byte device arrays
    for (long long n = 0; n < nn; ++n) {
      int es = 8;
      char *base = field_base(h, fi, (int)(lb[3] - 1 + n), &es);
      if (es == 4 && stage.empty()) {
        stage.resize((size_t)nbox);
        if (to_dev)
          for (long long m = 0; m < nbox; ++m) stage[m] = (float)host[m];
      }
      for (long long k = 0; k < nk; ++k) {
        char *dev = base + h->d.at(lb[0], lb[1], (int)(lb[2] + k)) * es;
        const long long hoff = ((n * nk + k) * ni) * nj;
        void *hp = es == 4 ? (void *)(stage.data() + hoff) : (void *)(host + hoff);
        if (to_dev)
          MLC_HIP(hipMemcpy2DAsync(dev, h->d.pj * es, hp, nj * es, nj * es, ni,
                                   hipMemcpyHostToDevice, h->stream));
        else
          MLC_HIP(hipMemcpy2DAsync(hp, nj * es, dev, h->d.pj * es, nj * es, ni,
                                   hipMemcpyDeviceToHost, h->stream));
      }
    }
    MLC_HIP(hipStreamSynchronize(h->stream));
    if (!to_dev && !stage.empty())
      for (long long m = 0; m < nbox; ++m) host[m] = (double)stage[m];
  }
  if (to_dev && fi.stat && !is_slice_field(f)) {
    h->have_static[f - MLC_FMZ] = true;
    h->inited = false;
    // keep host copies of the map factors for init_moloch's derived arrays
    std::vector<double> *hc = f == MLC_MX ? &h->h_mx
                              : f == MLC_MU ? &h->h_mu
                              : f == MLC_MV ? &h->h_mv : nullptr;
    if (hc) {
      for (long long i = 0; i < ni; ++i)
        for (long long j = 0; j < nj; ++j)
          (*hc)[h->d.at2((int)(lb[0] + j), (int)(lb[1] + i))] = host[i * nj + j];
    }
    // the column solve keeps double metrics in the mixed form
    if (h->mixed && (f == MLC_FMZ || f == MLC_FMZF)) {
      double *d64 = f == MLC_FMZ ? h->fmz64 : h->fmzf64;
      for (long long k = 0; k < nk; ++k) {
        double *dev = d64 + h->d.at(lb[0], lb[1], (int)(lb[2] + k));
        MLC_HIP(hipMemcpy2DAsync(dev, h->d.pj * sizeof(double),
                                 host + k * ni * nj, nj * sizeof(double),
                                 nj * sizeof(double), ni,
                                 hipMemcpyHostToDevice, h->stream));
      }
      MLC_HIP(hipStreamSynchronize(h->stream));
    }
  }
  if (to_dev && (f == MLC_TTEN || f == MLC_UTEN || f == MLC_VTEN || f == MLC_QXTEN ||
                 f == MLC_TKETEN || f == MLC_CHITEN))
    h->tend_dirty = true;
  return MLC_OK;
}

static void free_buf(Buf &b) {
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; }
}

static void free_all(mlc_handle *h) {
  Buf *bufs[] = {&h->u, &h->v, &h->t, &h->tvirt, &h->p, &h->rho, &h->qsat,
                 &h->ps, &h->tten, &h->uten, &h->vten, &h->qxten, &h->fmz,
                 &h->fmzf, &h->rfmzu, &h->rfmzv, &h->bdywtu, &h->bdywtv,
                 &h->bdywtw, &h->zeta, &h->dz, &h->mx, &h->mu, &h->mv, &h->hx,
                 &h->hy, &h->coru, &h->corv, &h->mx2, &h->rmx, &h->rmu,
                 &h->rmv, &h->tanrot, &h->s, &h->zdiv2, &h->zrfmu, &h->zrfmd,
                 &h->tke, &h->tketen, &h->chiten};
  for (auto b : bufs) free_buf(*b);
  for (auto &b : h->btab) free_buf(b);
  if (h->bidx) { (void)hipFree(h->bidx); h->bidx = nullptr; }
  for (int b = 0; b < 2; ++b)
    for (int s = 0; s < kMaxSlots; ++s) free_buf(h->adv[b][s]);
  for (int s = 0; s < kMaxSlots; ++s) free_buf(h->wz2[s]);
  if (!h->mixed) {       // aliases of fmz / fmzf, freed with them
    h->fmz64 = nullptr;
    h->fmzf64 = nullptr;
  }
  double **ds[] = {&h->w,     &h->tetavf, &h->colw, &h->vec1d, &h->red,
                   &h->fmz64, &h->fmzf64, &h->hgeo, &h->ztab};
  for (auto p : ds)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  for (auto &e : h->pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  h->pending.clear();
  h->pool.clear();
  if (h->pinned) { (void)hipHostFree(h->pinned); h->pinned = nullptr; h->pinned_len = 0; }
  {
    double **bp[] = {&h->lbc.ps0, &h->lbc.ps1, &h->lbc.u0, &h->lbc.u1, &h->lbc.v0,
                     &h->lbc.v1, &h->lbc.t0, &h->lbc.t1, &h->lbc.pai0, &h->lbc.pai1,
                     &h->lbc.qv0, &h->lbc.qv1, &h->lbc.qc0, &h->lbc.qc1, &h->lbc.qi0,
                     &h->lbc.qi1, &h->lbc.xfu, &h->lbc.xfv, &h->lbc.xfc, &h->lbc.tnudge,
                     &h->lbc.bvx, &h->lbc.bvy, &h->lbc.cnudge, &h->lbc.zn, &h->lbc.sx,
                     &h->lbc.sy};
    for (auto p : bp)
      if (*p) { (void)hipFree(*p); *p = nullptr; }
    h->lbc.have_spectral = false;
    h->lbc.on = false;
  }
  {
    double **sp[] = {&h->sl.zetaf, &h->sl.xlat, &h->sl.pf, &h->sl.th, &h->sl.rh,
                     &h->sl.omega, &h->sl.rhox, &h->sl.tp, &h->sl.th700, &h->sl.ptrop,
                     &h->sl.ktrop, &h->sl.kmxpbl};
    for (auto p : sp)
      if (*p) { (void)hipFree(*p); *p = nullptr; }
    h->sl.on = h->sl.have_zetaf = h->sl.have_xlat = false;
  }
  if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

// zero-filled like getmem (Share/mod_space.F90:309-310)
static int balloc(mlc_handle *h, Buf *b, long long n, int es) {
  b->es = es;
  MLC_HIP(hipMalloc(&b->p, (size_t)n * es));
  return launch_fill(h, *b, n, 0.0);
}

static int dalloc(mlc_handle *h, double **p, long long n) {
  MLC_HIP(hipMalloc(p, (size_t)n * sizeof(double)));
  return launch_fill64(h, *p, n, 0.0);
}

// Upload a host double array into a device Buf (converting to its type).
static int upload(mlc_handle *h, const Buf &b, const std::vector<double> &v) {
  if (b.es == 8) {
    MLC_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
    MLC_HIP(hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> f(v.begin(), v.end());
    MLC_HIP(hipMemcpyAsync(b.p, f.data(), f.size() * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
    MLC_HIP(hipStreamSynchronize(h->stream));
  }
  return MLC_OK;
}

}  // namespace mlc

namespace mlc {

// The boundary weights bdywtu, bdywtv, bdywtw as column ids and tables of
// their distinct columns (exact copies of the uploaded values).  Columns are
// compared byte for byte, so any upload is represented exactly; the weights
// setup_bdywt makes (Main/mod_bdycod.F90:4033-4047) have nspgx + 1 distinct
// columns at most.
static int build_bdy_tables(mlc_handle *h) {
  const long long P = h->d.plane;
  const int kz = h->d.kz;
  const size_t es = (size_t)h->es, colb = (size_t)kz * es;
  if (!h->bidx) MLC_HIP(hipMalloc(&h->bidx, (size_t)(3 * P) * sizeof(int)));
  std::vector<int> idx((size_t)(3 * P), 0);
  std::vector<unsigned char> host((size_t)P * colb), col(colb);
  const Buf *src[3] = {&h->bdywtu, &h->bdywtv, &h->bdywtw};
  for (int w = 0; w < 3; ++w) {
    MLC_HIP(hipMemcpyAsync(host.data(), src[w]->p, host.size(), hipMemcpyDeviceToHost,
                           h->stream));
    MLC_HIP(hipStreamSynchronize(h->stream));
    std::vector<unsigned char> tab;
    std::unordered_map<uint64_t, std::vector<int>> seen;
    int ncol = 0;
    for (long long c = 0; c < P; ++c) {
      uint64_t hsh = 1469598103934665603ull;       // FNV-1a over the column's bytes
      for (int k = 0; k < kz; ++k) {
        const unsigned char *p = host.data() + ((size_t)k * (size_t)P + (size_t)c) * es;
        std::memcpy(col.data() + (size_t)k * es, p, es);
        for (size_t b = 0; b < es; ++b) hsh = (hsh ^ p[b]) * 1099511628211ull;
      }
      int id = -1;
      auto &cands = seen[hsh];
      for (int cand : cands)
        if (std::memcmp(tab.data() + (size_t)cand * colb, col.data(), colb) == 0) {
          id = cand;
          break;
        }
      if (id < 0) {
        id = ncol++;
        cands.push_back(id);
        tab.insert(tab.end(), col.begin(), col.end());
      }
      idx[(size_t)w * (size_t)P + (size_t)c] = id;
    }
    free_buf(h->btab[w]);
    h->btab[w].es = (int)es;
    MLC_HIP(hipMalloc(&h->btab[w].p, tab.size()));
    MLC_HIP(hipMemcpyAsync(h->btab[w].p, tab.data(), tab.size(), hipMemcpyHostToDevice,
                           h->stream));
    MLC_HIP(hipStreamSynchronize(h->stream));
    h->bcols[w] = ncol;
  }
  MLC_HIP(hipMemcpyAsync(h->bidx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  MLC_HIP(hipStreamSynchronize(h->stream));
  return MLC_OK;
}

int create_handle(const mlc_grid_desc *g, hipStream_t shared, mlc_handle **out) {
  if (!g || !out) return fail(MLC_E_ARG, "null argument");
  *out = nullptr;
  if (g->jx < 6 || g->iy < 6 || g->kz < 4 || g->nadv < 1 || g->nsound < 1)
    return fail(MLC_E_ARG, "grid too small (jx, iy >= 6, kz >= 4) or nadv/nsound < 1");
  if (g->nqx < 1 || g->nqx > 10 || (g->ipptls > 1 && g->nqx < 5) ||
      (g->ipptls == 1 && g->nqx < 2) || g->ipptls < 0 || g->ipptls > 2)
    return fail(MLC_E_ARG, "nqx/ipptls combination not supported");
  if (g->precision != MLC_FP64 && g->precision != MLC_MIXED)
    return fail(MLC_E_ARG, "precision must be MLC_FP64 or MLC_MIXED");
  if (g->ntr < 0 || g->ntr > kMaxTr) return fail(MLC_E_ARG, "ntr must lie in 0..32");
  if (g->crm != 0 && g->crm != 1) return fail(MLC_E_ARG, "crm must be 0 or 1");
  if (g->crm && g->lrotllr)
    return fail(MLC_E_ARG, "a doubly periodic domain takes a normal projection, not ROTLLR");
  for (int r : g->reserved)
    if (r != 0) return fail(MLC_E_ARG, "reserved fields must be zero");
  if (!(g->dx > 0) || !(g->dtsec > 0) || !(g->mo_ztop > 0))
    return fail(MLC_E_ARG, "dx, dtsec and mo_ztop must be positive");
  mlc_handle *h = new (std::nothrow) mlc_handle();
  if (!h) return fail(MLC_E_ARG, "out of host memory");
  h->g = *g;
  if (hipGetDevice(&h->device) != hipSuccess) h->device = -1;
  h->mixed = g->precision == MLC_MIXED;
  h->es = h->mixed ? 4 : 8;
  // 8-row horizontal-WAF tiles for FP64 (fewer halo rows per output row),
  // 4 for the mixed form, whose float tiles leave room for more blocks: what
  // the autotuner chose at 256 x 256 x 41, 512 x 512 x 41 and 100 x 50 x 30
  h->tune.waf_tile = h->mixed ? 4 : 8;
  // the momentum update streams whole columns (fastest at 256x256x41 in
  // FP64, 20.3 against 21.8 us in the mixed form)
  h->tune.mom_chunk = g->kz;
  h->d = mlc::make_dims(g->jx, g->iy, g->kz, g->crm);
  // advected slots in the order of advection (Main/mod_moloch.F90:786-808):
  // tetav, pai, ux, vx, wx, qv, qx(iqfrst:nqx) with ipptls > 0, tkex, trac
  h->nslots = A_QX0 + g->nqx;
  for (int f = 0; f <= A_QX0; ++f) h->adv_list[h->nadvf++] = f;
  if (g->ipptls > 0)
    for (int n = 1; n < g->nqx; ++n) h->adv_list[h->nadvf++] = A_QX0 + n;
  if (g->ibltyp == 2) {
    h->s_tkex = h->nslots++;
    h->adv_list[h->nadvf++] = h->s_tkex;
  }
  if (g->ntr > 0) {
    h->s_tr0 = h->nslots;
    for (int n = 0; n < g->ntr; ++n) h->adv_list[h->nadvf++] = h->nslots++;
  }
  h->h_mx.assign((size_t)h->d.plane, 0.0);
  h->h_mu.assign((size_t)h->d.plane, 0.0);
  h->h_mv.assign((size_t)h->d.plane, 0.0);
  h->h_rlat.assign((size_t)g->iy, 0.0);
  h->h_ffilt.assign((size_t)g->kz, 0.0);
  int rc = MLC_OK;
  auto fin = [&](int code) {
    free_all(h);
    delete h;
    return code;
  };
  if (shared) {
    h->stream = shared;
    h->own_stream = false;
  } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    return fin(fail(MLC_E_HIP, "hipStreamCreate failed"));
  }
  const long long P = h->d.plane, kz = g->kz;
  const long long n3 = P * kz, n3p = P * (kz + 1);
  const int es = h->es;
  struct { Buf *b; long long n; } al[] = {
      {&h->u, n3}, {&h->v, n3}, {&h->t, n3}, {&h->tvirt, n3}, {&h->p, n3},
      {&h->rho, n3}, {&h->qsat, n3}, {&h->ps, P}, {&h->tten, n3},
      {&h->uten, n3}, {&h->vten, n3}, {&h->qxten, n3 * g->nqx},
      {&h->fmz, n3}, {&h->fmzf, n3p}, {&h->rfmzu, n3}, {&h->rfmzv, n3},
      {&h->bdywtu, n3}, {&h->bdywtv, n3}, {&h->bdywtw, n3}, {&h->zeta, n3},
      {&h->dz, n3}, {&h->mx, P}, {&h->mu, P}, {&h->mv, P}, {&h->hx, P},
      {&h->hy, P}, {&h->coru, P}, {&h->corv, P}, {&h->mx2, P}, {&h->rmx, P},
      {&h->rmu, P}, {&h->rmv, P}, {&h->tanrot, P}, {&h->s, n3p},
      {&h->zdiv2, n3}, {&h->zrfmu, n3}, {&h->zrfmd, n3}};
  for (auto &a : al)
    if ((rc = balloc(h, a.b, a.n, es))) return fin(rc);
  if ((rc = dalloc(h, &h->w, n3p))) return fin(rc);
  if ((rc = dalloc(h, &h->tetavf, n3))) return fin(rc);
  if ((rc = dalloc(h, &h->colw, 3 * n3))) return fin(rc);
  if ((rc = dalloc(h, &h->vec1d, 5 * kz + 1))) return fin(rc);
  if ((rc = dalloc(h, &h->hgeo, 3 * P))) return fin(rc);
  if ((rc = dalloc(h, &h->ztab, 2 * (2 * kz + 1)))) return fin(rc);
  for (int m = 0; m < 3; ++m) h->geo[m] = h->hgeo + m * P;
  if (h->mixed) {
    if ((rc = dalloc(h, &h->fmz64, n3))) return fin(rc);
    if ((rc = dalloc(h, &h->fmzf64, n3p))) return fin(rc);
  } else {
    h->fmz64 = h->fmz.as<double>();
    h->fmzf64 = h->fmzf.as<double>();
  }
  for (int s = 0; s < h->nslots; ++s)
    for (int b = 0; b < 2; ++b)
      if ((rc = balloc(h, &h->adv[b][s], n3, s == A_PAI || s == A_TETAV ? 8 : es)))
        return fin(rc);
  for (int m = 0; m < h->nadvf; ++m) {
    const int s = h->adv_list[m];
    if ((rc = balloc(h, &h->wz2[s], n3, s == A_PAI || s == A_TETAV ? 8 : es)))
      return fin(rc);
  }
  if (g->ibltyp == 2 && ((rc = balloc(h, &h->tke, n3p, es)) ||
                         (rc = balloc(h, &h->tketen, n3p, es))))
    return fin(rc);
  if (g->ntr > 0 && (rc = balloc(h, &h->chiten, n3 * g->ntr, es))) return fin(rc);
  const long long nblk = ((h->d.je2 + 63) / 64) * ((h->d.ie2 + 3) / 4) * kz;
  h->red_len = (size_t)(nblk * 8 + 64);
  if ((rc = dalloc(h, &h->red, (long long)h->red_len))) return fin(rc);
  if (hipStreamSynchronize(h->stream) != hipSuccess)
    return fin(fail(MLC_E_HIP, "synchronize after allocation failed"));
  *out = h;
  return MLC_OK;
}

void destroy_handle(mlc_handle *h) {
  free_all(h);
  delete h;
}

// nadv x (mo_nsound acoustic sub-steps + advection), without the tail
static int dycore_passes(mlc_handle *h) {
  int rc;
  for (int nadv = 1; nadv <= h->g.nadv; ++nadv) {
    if ((rc = launch_sound(h))) return rc;
    if ((rc = launch_advection(h))) return rc;
  }
  return MLC_OK;
}

// One physics-off step after the tendency reset: the tendencies are zero.
int step_body(mlc_handle *h) {
  int rc;
  if ((rc = dycore_passes(h))) return rc;
  return launch_step_tail(h);
}

}  // namespace mlc

using namespace mlc;

extern "C" {

const char *mlc_last_error(void) { return g_err.c_str(); }

int mlc_create(const mlc_grid_desc *g, mlc_handle **out) {
  return create_handle(g, nullptr, out);
}

int mlc_destroy(mlc_handle *h) {
  if (!h) return MLC_OK;
  if (h->batch_member)
    return fail(MLC_E_ARG, "handle belongs to a batch: destroy the batch");
  destroy_handle(h);
  return MLC_OK;
}

int mlc_set_static(mlc_handle *h, int field, const double *host,
                   const int lb[4], const int ub[4]) {
  FieldInfo fi;
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!field_info(h, field, &fi) || !fi.stat)
    return fail(MLC_E_ARG, "not a static field id");
  return copy_box(h, field, const_cast<double *>(host), lb, ub, true);
}

int mlc_put_state(mlc_handle *h, int field, const double *host,
                  const int lb[4], const int ub[4]) {
  FieldInfo fi;
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!field_info(h, field, &fi) || fi.stat)
    return fail(MLC_E_ARG, "not a state field id");
  return copy_box(h, field, const_cast<double *>(host), lb, ub, true);
}

int mlc_get_state(mlc_handle *h, int field, double *host, const int lb[4],
                  const int ub[4]) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  return copy_box(h, field, host, lb, ub, false);
}

int mlc_init(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  const int need[] = {MLC_FMZ, MLC_FMZF, MLC_RFMZU, MLC_RFMZV, MLC_BDYWTU,
                      MLC_BDYWTV, MLC_BDYWTW, MLC_ZETA, MLC_MX, MLC_MU,
                      MLC_MV, MLC_HX, MLC_HY, MLC_CORU, MLC_CORV, MLC_FFILT};
  for (int f : need)
    if (!h->have_static[f - MLC_FMZ]) {
      char buf[64];
      snprintf(buf, sizeof buf, "static field %d not set before mlc_init", f);
      return fail(MLC_E_STATE, buf);
    }
  if (h->g.lrotllr && !h->have_static[MLC_RLAT - MLC_FMZ])
    return fail(MLC_E_STATE, "ROTLLR needs MLC_RLAT");
  const int jx = h->d.jx, iy = h->d.iy, kz = h->d.kz;
  const long long P = h->d.plane;
  // init_moloch, Main/mod_moloch.F90:263-268
  std::vector<double> mx2(P, 0.0), rmx(P, 0.0), rmu(P, 0.0), rmv(P, 0.0),
      tanrot(P, 0.0);
  for (int i = 1; i <= iy; ++i)
    for (int j = 1; j <= jx; ++j) {
      const long long c = h->d.at2(j, i);
      mx2[c] = h->h_mx[c] * h->h_mx[c];
      rmx[c] = 1.0 / h->h_mx[c];
      rmu[c] = 1.0 / h->h_mu[c];
      rmv[c] = 1.0 / h->h_mv[c];
    }
  // ROTLLR curvature factor of :813-814, hoisted out of the level loop
  if (h->g.lrotllr)
    for (int i = 2; i <= iy - 2; ++i)
      for (int j = 2; j <= jx - 2; ++j) {
        const long long c = h->d.at2(j, i);
        const double dlat = kDegrad * 0.5 * (h->h_rlat[i - 1] + h->h_rlat[i]);
        tanrot[c] = std::sin(dlat) * h->h_mx[c] * kRearthrad;
      }
  // vertical grid (Main/mod_params.F90:2460-2463) and :273-275, :294-299
  std::vector<double> zf(kz + 1), zh(kz);
  const double ztop = h->g.mo_ztop, a0 = h->g.mo_a0;
  {
    const double dz = ztop / (double)kz;
    zf[kz] = 0.0;
    zf[0] = ztop;
    for (int k = kz; k >= 2; --k) zf[k - 1] = zf[k] + dz;
    zh[kz - 1] = dz * 0.5;
    zh[0] = ztop - dz * 0.5;
    for (int k = kz - 1; k >= 2; --k) zh[k - 1] = zh[k] + dz;
  }
  std::vector<double> v1(5 * kz + 1);
  for (int k = 1; k <= kz + 1; ++k) v1[k - 1] = gzita(zf[k - 1], ztop, a0);
  for (int k = 1; k <= kz; ++k) {
    v1[kz + 1 + k - 1] = gzita(zh[k - 1], ztop, a0);
    v1[2 * kz + 1 + k - 1] =
        0.125 * 0.850 * (1.0 / ((double)k + 1.0) - 1.0 / ((double)kz + 2.0));
    v1[3 * kz + 1 + k - 1] =
        0.125 * (0.55 + 0.45 * ((double)(kz - k + 1) - 1.0) / ((double)kz - 1.0));
    v1[4 * kz + 1 + k - 1] = h->h_ffilt[k - 1];
  }
  const double mo_dzita = zf[kz - 1];
  h->rdzita = 1.0 / mo_dzita;
  h->rdx = 1.0 / h->g.dx;
  h->dtstepa = h->g.dtsec / (double)h->g.nadv;
  h->dtsound = h->dtstepa / (double)h->g.nsound;
  int rc;
  if ((rc = upload(h, h->mx2, mx2))) return rc;
  if ((rc = upload(h, h->rmx, rmx))) return rc;
  if ((rc = upload(h, h->rmu, rmu))) return rc;
  if ((rc = upload(h, h->rmv, rmv))) return rc;
  if ((rc = upload(h, h->tanrot, tanrot))) return rc;
  MLC_HIP(hipMemcpyAsync(h->vec1d, v1.data(), v1.size() * sizeof(double),
                         hipMemcpyHostToDevice, h->stream));
  // w(:,:,1) = 0  (:276-278)
  if ((rc = launch_fill64(h, h->w, P, 0.0))) return rc;
  // wafone's vertical metric ratios (:888-889), fixed for the run
  if ((rc = launch_waf_metrics(h))) return rc;
  if ((rc = build_bdy_tables(h))) return rc;
  // the terrain metrics formed from the orography, where they reproduce
  // every uploaded value (md_fmz: gzitap(z)*(ht*regrav) + bzitap(z))
  h->zmet_ok = false;
  h->tune.terrain = 0;
  if (h->mo_h > 0.0 && h->have_static[MLC_HT - MLC_FMZ] &&
      h->have_static[MLC_HTU - MLC_FMZ] && h->have_static[MLC_HTV - MLC_FMZ]) {
    const int nt = 2 * kz + 1;
    std::vector<double> zt(2 * (size_t)nt);
    for (int k = 1; k <= kz; ++k) {
      zt[k - 1] = gzitap(zh[k - 1], ztop, a0);
      zt[nt + k - 1] = bzitap(zh[k - 1], ztop, h->mo_h);
    }
    for (int k = 1; k <= kz + 1; ++k) {
      zt[kz + k - 1] = gzitap(zf[k - 1], ztop, a0);
      zt[nt + kz + k - 1] = bzitap(zf[k - 1], ztop, h->mo_h);
    }
    MLC_HIP(hipMemcpyAsync(h->ztab, zt.data(), zt.size() * sizeof(double),
                           hipMemcpyHostToDevice, h->stream));
    long long nbad = -1;
    if ((rc = launch_terrain_check(h, &nbad))) return rc;
    h->zmet_ok = nbad == 0;
    // default stages (measured at 256 x 256 x 41): in FP64 the divergence,
    // the column solve and the staggering gain, in the mixed form the column
    // solve and the staggering; the horizontal WAF's tile setup and the
    // float divergence lose to their extra divisions
    h->tune.terrain = h->zmet_ok ? (h->mixed ? 6 : 7) : 0;
    ++h->config_gen;
  }
  MLC_HIP(hipStreamSynchronize(h->stream));
  h->inited = true;
  return MLC_OK;
}

int mlc_reset_tendencies(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  const long long n3 = h->d.plane * h->d.kz;
  int rc;
  if ((rc = launch_fill(h, h->tten, n3, 0.0))) return rc;
  if ((rc = launch_fill(h, h->uten, n3, 0.0))) return rc;
  if ((rc = launch_fill(h, h->vten, n3, 0.0))) return rc;
  if ((rc = launch_fill(h, h->qxten, n3 * h->g.nqx, 0.0))) return rc;
  if (h->s_tkex >= 0 && (rc = launch_fill(h, h->tketen, n3 + h->d.plane, 0.0))) return rc;
  if (h->s_tr0 >= 0 && (rc = launch_fill(h, h->chiten, n3 * h->g.ntr, 0.0))) return rc;
  h->tend_dirty = false;
  return MLC_OK;
}

int mlc_dynamical_core(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  const int rc = dycore_passes(h);
  return rc ? rc : launch_dyn_tail(h);
}

int mlc_diagnostics(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  return launch_diagnostics(h);
}

int mlc_status_update(mlc_handle *h, double dtinc) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  return launch_status_update(h, dtinc);
}

int mlc_set_param(mlc_handle *h, int param, double value) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  switch (param) {
    case MLC_PARAM_TKEMIN:
      if (!std::isfinite(value)) return fail(MLC_E_ARG, "tkemin must be finite");
      h->tkemin = value;
      ++h->config_gen;
      return MLC_OK;
    case MLC_PARAM_MO_H:
      if (!std::isfinite(value) || value <= 0.0) return fail(MLC_E_ARG, "mo_h must be > 0");
      h->mo_h = value;
      h->inited = false;
      return MLC_OK;
    default: return fail(MLC_E_ARG, "unknown MLC_PARAM_*");
  }
}

int mlc_step(mlc_handle *h, int nsteps) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  if (nsteps < 0) return fail(MLC_E_ARG, "nsteps < 0");
  int rc;
  for (int n = 0; n < nsteps; ++n) {
    // tendencies are only non-zero after a caller uploaded some; zeroing
    // zeros is skipped
    if (h->tend_dirty && (rc = mlc_reset_tendencies(h))) return rc;
    if ((rc = step_body(h))) return rc;
  }
  return MLC_OK;
}

int mlc_lbc_configure(mlc_handle *h, const mlc_lbc_desc *desc) {
  if (!h || !desc) return fail(MLC_E_ARG, "null argument");
  if (desc->reserved != 0) return fail(MLC_E_ARG, "reserved fields must be zero");
  if (h->d.crm)   // do_apply_bdy is false there (Main/mod_moloch.F90:305, irceideal)
    return fail(MLC_E_ARG, "a doubly periodic domain has no lateral boundary");
  if (desc->top_nudge && (desc->nztop < 0 || desc->nztop > h->g.kz))
    return fail(MLC_E_ARG, "nztop must lie in 0..kz");
  if (desc->spectral_nudge &&
      (desc->km < 1 || desc->lm < 1 || desc->km > 64 || desc->lm > 64 || !(desc->dtrad > 0)))
    return fail(MLC_E_ARG, "spectral nudging needs 1 <= km, lm <= 64 and dtrad > 0");
  if (h->lbc.on && desc->spectral_nudge &&
      (desc->km != h->lbc.desc.km || desc->lm != h->lbc.desc.lm))
    h->lbc.have_spectral = false;            // the bases must be set again
  h->lbc.desc = *desc;
  if (h->lbc.on) return MLC_OK;
  const long long P = h->d.plane, n3 = P * h->g.kz;
  double **b3[] = {&h->lbc.u0, &h->lbc.u1, &h->lbc.v0, &h->lbc.v1, &h->lbc.t0, &h->lbc.t1,
                   &h->lbc.pai0, &h->lbc.pai1, &h->lbc.qv0, &h->lbc.qv1, &h->lbc.qc0,
                   &h->lbc.qc1, &h->lbc.qi0, &h->lbc.qi1};
  int rc;
  for (auto p : b3)
    if ((rc = dalloc(h, p, n3))) return rc;
  if ((rc = dalloc(h, &h->lbc.ps0, P)) || (rc = dalloc(h, &h->lbc.ps1, P))) return rc;
  double **wts[] = {&h->lbc.xfu, &h->lbc.xfv, &h->lbc.xfc};
  for (auto p : wts) {
    MLC_HIP(hipMalloc(p, (size_t)n3 * sizeof(double)));
    if ((rc = launch_fill64(h, *p, n3, -1.0))) return rc;     // no relaxation zone yet
  }
  if ((rc = dalloc(h, &h->lbc.tnudge, h->g.kz))) return rc;
  MLC_HIP(hipStreamSynchronize(h->stream));
  h->lbc.on = true;
  return MLC_OK;
}

int mlc_lbc_spectral(mlc_handle *h, const double *bvx, const double *bvy,
                     const double *cnudge) {
  if (!h || !bvx || !bvy || !cnudge) return fail(MLC_E_ARG, "null argument");
  if (!h->lbc.on || !h->lbc.desc.spectral_nudge)
    return fail(MLC_E_STATE, "configure spectral nudging with mlc_lbc_configure first");
  const long long km2 = 2LL * h->lbc.desc.km, lm2 = 2LL * h->lbc.desc.lm;
  const long long jx = h->g.jx, iy = h->g.iy, kz = h->g.kz;
  double **bp[] = {&h->lbc.bvx, &h->lbc.bvy, &h->lbc.cnudge, &h->lbc.zn, &h->lbc.sx,
                   &h->lbc.sy};
  for (auto p : bp)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  int rc;
  if ((rc = dalloc(h, &h->lbc.bvx, jx * km2)) || (rc = dalloc(h, &h->lbc.bvy, iy * lm2)) ||
      (rc = dalloc(h, &h->lbc.cnudge, kz)) || (rc = dalloc(h, &h->lbc.zn, h->d.plane * kz)) ||
      (rc = dalloc(h, &h->lbc.sx, iy * km2 * kz)) || (rc = dalloc(h, &h->lbc.sy, jx * lm2 * kz)))
    return rc;
  MLC_HIP(hipMemcpy(h->lbc.bvx, bvx, (size_t)(jx * km2) * sizeof(double), hipMemcpyHostToDevice));
  MLC_HIP(hipMemcpy(h->lbc.bvy, bvy, (size_t)(iy * lm2) * sizeof(double), hipMemcpyHostToDevice));
  MLC_HIP(hipMemcpy(h->lbc.cnudge, cnudge, (size_t)kz * sizeof(double), hipMemcpyHostToDevice));
  h->lbc.have_spectral = true;
  h->lbc.tspectral = 0.0;
  return MLC_OK;
}

int mlc_boundary(mlc_handle *h, double x1) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  if (!h->lbc.on) return fail(MLC_E_STATE, "mlc_lbc_configure has not been called");
  if (h->lbc.desc.spectral_nudge && !h->lbc.have_spectral)
    return fail(MLC_E_STATE, "mlc_lbc_spectral has not been called");
  return launch_boundary(h, x1);
}

int mlc_step_lbc(mlc_handle *h, int nsteps, const double *x1) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  if (!h->lbc.on) return fail(MLC_E_STATE, "mlc_lbc_configure has not been called");
  if (nsteps < 0 || (nsteps > 0 && !x1)) return fail(MLC_E_ARG, "bad nsteps or x1");
  if (h->lbc.desc.spectral_nudge && !h->lbc.have_spectral)
    return fail(MLC_E_STATE, "mlc_lbc_spectral has not been called");
  if (h->g.ntr > 0)
    return fail(MLC_E_ARG, "with tracers the chemistry boundary relaxation (morelax_chiten) "
                           "is the caller's: step with mlc_dynamical_core, mlc_boundary, "
                           "mlc_diagnostics and mlc_status_update");
  int rc;
  for (int n = 0; n < nsteps; ++n) {
    if (h->tend_dirty && (rc = mlc_reset_tendencies(h))) return rc;
    if ((rc = mlc_dynamical_core(h))) return rc;
    if ((rc = launch_boundary(h, x1[n]))) return rc;
    if ((rc = launch_diagnostics(h))) return rc;
    if ((rc = launch_status_update(h, h->g.dtsec))) return rc;
  }
  return MLC_OK;
}

int mlc_mkslice(mlc_handle *h, const mlc_slice_desc *desc) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!desc) return fail(MLC_E_ARG, "null slice descriptor");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  if (!h->sl.have_zetaf) return fail(MLC_E_STATE, "MLC_ZETAF has not been set");
  if (desc->irceideal != 1 && !h->sl.have_xlat)
    return fail(MLC_E_STATE, "MLC_XLAT has not been set (needed unless irceideal = 1)");
  if (desc->irceideal != 1 && !(desc->dayspy > 0.0))
    return fail(MLC_E_ARG, "dayspy must be positive");
  if (!(desc->rhmin <= desc->rhmax)) return fail(MLC_E_ARG, "rhmin > rhmax");
  if (h->d.jx < 4 || h->d.iy < 4) return fail(MLC_E_ARG, "domain too small for mkslice");
  return launch_mkslice(h, *desc);
}

int mlc_init_restart(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  if (!h->inited) return fail(MLC_E_STATE, "mlc_init has not been called");
  return launch_init_restart(h);
}

int mlc_massck(mlc_handle *h, double out[4]) {
  if (!h || !out) return fail(MLC_E_ARG, "null argument");
  return launch_massck(h, out);
}

int mlc_massck_noalloc(mlc_handle *h, double out[4]) {
  if (!h || !out) return fail(MLC_E_ARG, "null argument");
  return launch_massck_noalloc(h, out);
}

int mlc_check_finite(mlc_handle *h, double *psmax, double *psmin) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  return launch_check_finite(h, psmax, psmin);
}

static int set_variant(mlc_handle *h, int stage, int value);

int mlc_set_variant(mlc_handle *h, int stage, int value) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  const int rc = set_variant(h, stage, value);
  if (rc == MLC_OK) ++h->config_gen;   // recorded batch plans copied the old shapes
  return rc;
}

static int set_variant(mlc_handle *h, int stage, int value) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  switch (stage) {
    case MLC_VAR_WAF_H:
      if (value < 0 || value > 1) return fail(MLC_E_ARG, "MLC_VAR_WAF_H takes 0 or 1");
      h->tune.waf_h = value;
      return MLC_OK;
    case MLC_VAR_WAF_ROWS:
      if (value < 4 || value > 256) return fail(MLC_E_ARG, "MLC_VAR_WAF_ROWS takes 4..256");
      h->tune.waf_rows = value;
      return MLC_OK;
    case MLC_VAR_WAF_GROUP:
      if (value != 2 && value != 4) return fail(MLC_E_ARG, "MLC_VAR_WAF_GROUP takes 2 or 4");
      h->tune.waf_group = value;
      return MLC_OK;
    case MLC_VAR_LEVEL_CHUNK:
      if (value < 1 || value > h->g.kz) return fail(MLC_E_ARG, "MLC_VAR_LEVEL_CHUNK takes 1..kz");
      h->tune.level_chunk = value;
      return MLC_OK;
    case MLC_VAR_MOM_CHUNK:
      if (value < 0 || value > h->g.kz) return fail(MLC_E_ARG, "MLC_VAR_MOM_CHUNK takes 0..kz");
      h->tune.mom_chunk = value;
      return MLC_OK;
    case MLC_VAR_STAGGER_CHUNK:
      if (value < 1 || value > h->g.kz) return fail(MLC_E_ARG, "MLC_VAR_STAGGER_CHUNK takes 1..kz");
      h->tune.stag_chunk = value;
      return MLC_OK;
    case MLC_VAR_COLUMN:
      if (value < 0 || value > 2) return fail(MLC_E_ARG, "MLC_VAR_COLUMN takes 0, 1 or 2");
      h->tune.column = value;
      return MLC_OK;
    case MLC_VAR_WAFV_GROUP:
      if (value != 1 && value != 2 && value != 4)
        return fail(MLC_E_ARG, "MLC_VAR_WAFV_GROUP takes 1, 2 or 4");
      h->tune.wafv_group = value;
      return MLC_OK;
    case MLC_VAR_XCD_MAP:
      if (value < 0 || value > 1) return fail(MLC_E_ARG, "MLC_VAR_XCD_MAP takes 0 or 1");
      h->tune.xcd_map = value;
      return MLC_OK;
    case MLC_VAR_WAF_TILE:
      if (value != 4 && value != 8) return fail(MLC_E_ARG, "MLC_VAR_WAF_TILE takes 4 or 8");
      h->tune.waf_tile = value;
      return MLC_OK;
    case MLC_VAR_WAF_HGROUP:
      if (value < 0 || value > kMaxAdv)
        return fail(MLC_E_ARG, "MLC_VAR_WAF_HGROUP takes 0 (all) .. 16");
      h->tune.waf_hgroup = value;
      return MLC_OK;
    case MLC_VAR_COLUMN_J:
      if (value != 16 && value != 32 && value != 64)
        return fail(MLC_E_ARG, "MLC_VAR_COLUMN_J takes 16, 32 or 64");
      h->tune.column_j = value;
      return MLC_OK;
    case MLC_VAR_MIXED_MERGE:
      if (value < 0 || value > 7) return fail(MLC_E_ARG, "MLC_VAR_MIXED_MERGE takes 0..7");
      h->tune.mixed_merge = value;
      return MLC_OK;
    case MLC_VAR_TERRAIN:
      if (value < 0 || value > 15) return fail(MLC_E_ARG, "MLC_VAR_TERRAIN takes 0..15");
      if (value != 0 && !h->zmet_ok)
        return fail(MLC_E_STATE,
                    "MLC_VAR_TERRAIN != 0 needs mlc_init to have verified the formed metrics "
                    "(MLC_PARAM_MO_H, MLC_HT, MLC_HTU, MLC_HTV)");
      h->tune.terrain = value;
      return MLC_OK;
    default:
      return fail(MLC_E_ARG, "unknown stage");
  }
}

int mlc_get_variant(const mlc_handle *h, int stage) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  switch (stage) {
    case MLC_VAR_WAF_H: return h->tune.waf_h;
    case MLC_VAR_WAF_ROWS: return h->tune.waf_rows;
    case MLC_VAR_WAF_GROUP: return h->tune.waf_group;
    case MLC_VAR_LEVEL_CHUNK: return h->tune.level_chunk;
    case MLC_VAR_COLUMN: return h->tune.column;
    case MLC_VAR_WAFV_GROUP: return h->tune.wafv_group;
    case MLC_VAR_XCD_MAP: return h->tune.xcd_map;
    case MLC_VAR_WAF_TILE: return h->tune.waf_tile;
    case MLC_VAR_WAF_HGROUP: return h->tune.waf_hgroup;
    case MLC_VAR_COLUMN_J: return h->tune.column_j;
    case MLC_VAR_MIXED_MERGE: return h->tune.mixed_merge;
    case MLC_VAR_TERRAIN: return h->tune.terrain;
    case MLC_VAR_MOM_CHUNK: return h->tune.mom_chunk;
    case MLC_VAR_STAGGER_CHUNK: return h->tune.stag_chunk;
    default: return fail(MLC_E_ARG, "unknown stage");
  }
}

int mlc_synchronize(mlc_handle *h) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  MLC_HIP(hipStreamSynchronize(h->stream));
  return MLC_OK;
}

int mlc_profile(mlc_handle *h, int enable) {
  if (!h) return fail(MLC_E_ARG, "null handle");
  int rc = prof_resolve(h);
  if (rc) return rc;
  for (int s = 0; s < K_COUNT; ++s) {
    h->launches[s] = 0;
    h->ms[s] = 0.0;
  }
  h->prof = enable != 0;
  // enable: 0 off, 1 every slot, otherwise bit (slot + 1) selects a slot
  h->prof_mask = enable == 1 ? ~0u : ((unsigned)enable >> 1);
  return MLC_OK;
}

int mlc_kernel_count(void) { return K_COUNT; }

const char *mlc_kernel_name(int slot) {
  return (slot >= 0 && slot < K_COUNT) ? kKernelNames[slot] : nullptr;
}

int mlc_kernel_stats(mlc_handle *h, int slot, long long *launches,
                     double *total_ms) {
  if (!h || slot < 0 || slot >= K_COUNT) return fail(MLC_E_ARG, "bad slot");
  int rc = prof_resolve(h);
  if (rc) return rc;
  if (launches) *launches = h->launches[slot];
  if (total_ms) *total_ms = h->ms[slot];
  return MLC_OK;
}

long long mlc_npoints(const mlc_handle *h) {
  if (!h) return 0;
  return (long long)h->d.je2 * h->d.ie2 * h->d.kz;
}

double mlc_constant(const char *name) {
  if (!name) return NAN;
  struct { const char *n; double v; } tab[] = {
      {"egrav", kEgrav}, {"rgas", kRgas},   {"cpd", kCpd},     {"cvd", kCvd},
      {"rdrcv", kRdrcv}, {"cpovr", kCpovr}, {"govr", kGovr},   {"govcp", kGovcp},
      {"ep1", kEp1},     {"ep2", kEp2},     {"tzero", kTzero}, {"p00", kP00},
      {"lrate", kLrate}, {"degrad", kDegrad}, {"mathpi", kMathpi},
      {"earthrad", kEarthrad}, {"eomeg2", kEomeg2}};
  for (auto &e : tab)
    if (std::strcmp(e.n, name) == 0) return e.v;
  return NAN;
}

}  // extern "C"
