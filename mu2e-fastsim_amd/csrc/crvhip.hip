// libcrvhip: context, tables, workspace, the C ABI of include/crvhip.h and the chained FP64 run.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <map>
#include <set>

#include "crv_internal.h"

struct crv_ctx {
  crv::Ctx c;
};

namespace crv {

static std::string g_create_error;

int set_error(Ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

void *ws(Ctx *c, int slot, size_t bytes) {
  if ((int)c->ws.size() <= slot) {
    c->ws.resize(slot + 1, nullptr);
    c->ws_size.resize(slot + 1, 0);
  }
  if (bytes == 0) bytes = 8;
  if (c->ws_size[slot] < bytes) {
    if (c->ws[slot]) (void)hipFree(c->ws[slot]);
    c->ws[slot] = nullptr;
    c->ws_size[slot] = 0;
    size_t b = bytes + bytes / 8;
    if (hipMalloc(&c->ws[slot], b) != hipSuccess) {
      (void)hipGetLastError();
      c->ws[slot] = nullptr;
      bool freed = false;
      for (const int sp : c->ws_spare)
        if (sp != slot && sp < (int)c->ws.size() && c->ws[sp]) {
          if (!freed) (void)hipDeviceSynchronize();  // no kernel still reads a spare slot
          (void)hipFree(c->ws[sp]);
          c->ws[sp] = nullptr;
          c->ws_size[sp] = 0;
          freed = true;
        }
      c->ws_spare.clear();
      if (!freed || hipMalloc(&c->ws[slot], b) != hipSuccess) {
        (void)hipGetLastError();
        c->ws[slot] = nullptr;
        c->ws_fail_slot = slot;
        c->ws_fail_bytes = b;
        return nullptr;
      }
    }
    c->ws_size[slot] = b;
  }
  return c->ws[slot];
}

int ws_fail(Ctx *c) {
  return set_error(c, CRV_E_HIP, "workspace allocation failed (slot " + std::to_string(c->ws_fail_slot) + ", " +
                                     std::to_string(c->ws_fail_bytes >> 20) + " MiB)");
}

namespace {

__global__ void k_i32_to_i64(const int32_t *__restrict__ in, int64_t *__restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

template <class T>
int upload(Ctx *c, const T *host, int64_t n, const T **dst) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<int64_t>(n, 1) * sizeof(T)) != hipSuccess) return set_error(c, CRV_E_HIP, "hipMalloc failed");
  c->owned.push_back(p);
  if (n && hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return set_error(c, CRV_E_HIP, "hipMemcpy failed");
  *dst = (const T *)p;
  return CRV_OK;
}

}  // namespace

int scan_i64(Ctx *c, const int64_t *in, int64_t *out, int64_t n, int slot, hipStream_t s) {
  if (slot < 0 || slot >= WS_STAGE_SLOTS) return set_error(c, CRV_E_ARG, "scan: workspace slot out of range");
  CRV_HIP(c, hipMemsetAsync(out, 0, 8, s));
  if (n == 0) return CRV_OK;
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out + 1, (int)n, s));
  void *tmp = ws(c, WS_SCAN_TMP + slot, tb);
  if (!tmp) return set_error(c, CRV_E_HIP, "scan workspace allocation failed");
  CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out + 1, (int)n, s));
  return CRV_OK;
}

int scan_i32_to_i64(Ctx *c, const int32_t *in, int64_t *out, int64_t n, int slot, hipStream_t s) {
  int64_t *w = (int64_t *)ws(c, WS_SCAN_WIDE + slot, (n + 1) * 8);
  if (!w) return set_error(c, CRV_E_HIP, "scan workspace allocation failed");
  if (n) k_i32_to_i64<<<nblocks(n, 256), 256, 0, s>>>(in, w, n);
  CRV_HIP(c, hipGetLastError());
  return scan_i64(c, w, out, n, slot, s);
}

int read_i64(Ctx *c, const int64_t *dev, int64_t *host, hipStream_t s) {
  CRV_HIP(c, hipMemcpyAsync(host, dev, 8, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

// ---- reco waveforms: continuing digis of one channel, concatenated (reco/pulsefit.py:319-345) ----
namespace {

__global__ void k_wf_flags(int64_t n, int Sz, const int64_t *__restrict__ label, const int64_t *__restrict__ tdc,
                           uint8_t *__restrict__ flag, int *__restrict__ flags) {
  int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  bool nw = d == 0 || label[d] != label[d - 1] || tdc[d] != tdc[d - 1] + Sz;
  if (d > 0 && (label[d] < label[d - 1] || (label[d] == label[d - 1] && tdc[d] < tdc[d - 1])))
    atomicAdd(&flags[F_UNSORTED], 1);
  flag[d] = nw;
}

__global__ void k_wf_table(int64_t nw, int64_t nd, int Sz, const int32_t *__restrict__ first,
                           const int64_t *__restrict__ label, const int64_t *__restrict__ tdc, crv_waveforms o) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  int64_t f = first[w], e = w + 1 < nw ? first[w + 1] : nd;
  int64_t lab = label[f];
  o.ev[w] = lab / CRV_CHAN_MULT;
  o.chan[w] = (int32_t)(lab % CRV_CHAN_MULT);
  o.start[w] = tdc[f];
  o.nsamp[w] = (int32_t)((e - f) * Sz);
  o.off[w] = f * Sz;
  if (w == nw - 1) o.off[nw] = nd * Sz;
}

__global__ void k_copy_i16(const int16_t *__restrict__ in, int16_t *__restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

}  // namespace

int waveforms_stage(Ctx *c, const int64_t *label, const int64_t *starttdc, const int16_t *adc, int64_t n,
                    crv_waveforms *out, hipStream_t s) {
  const int Sz = c->p.samples_zs;
  if (n == 0) {
    out->n = 0;
    out->n_samples = 0;
    return CRV_OK;
  }
  uint8_t *flag = (uint8_t *)ws(c, 130, n + 8);
  int32_t *first = (int32_t *)ws(c, 131, (n + 1) * 4);
  int32_t *nsel = (int32_t *)ws(c, 132, 8);
  if (!flag || !first || !nsel) return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  k_wf_flags<<<nblocks(n, 256), 256, 0, s>>>(n, Sz, label, starttdc, flag, c->flags);
  CRV_HIP(c, hipGetLastError());
  size_t tb = 0;
  hipcub::CountingInputIterator<int32_t> it(0);
  CRV_HIP(c, hipcub::DeviceSelect::Flagged(nullptr, tb, it, flag, first, nsel, (int)n, s));
  void *tmp = ws(c, 133, tb);
  CRV_HIP(c, hipcub::DeviceSelect::Flagged(tmp, tb, it, flag, first, nsel, (int)n, s));
  int32_t hw = 0;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(&hw, nsel, 4, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_UNSORTED]) return set_error(c, CRV_E_ARG, "digis must be sorted by (label, start TDC)");
  out->n = hw;
  out->n_samples = n * Sz;
  if (out->capacity < 0) {  // internal outputs of the chained run
    out->ev = (int64_t *)ws(c, 134, hw * 8);
    out->chan = (int32_t *)ws(c, 135, hw * 4);
    out->start = (int64_t *)ws(c, 136, hw * 8);
    out->nsamp = (int32_t *)ws(c, 137, hw * 4);
    out->off = (int64_t *)ws(c, 138, (hw + 1) * 8);
    out->adc = nullptr;  // the digi samples are the waveform samples
    out->capacity = hw;
    out->sample_capacity = n * Sz;
  }
  if (hw > out->capacity || (out->adc && n * Sz > out->sample_capacity))
    return set_error(c, CRV_E_CAPACITY, "reco waveforms: output capacity too small");
  k_wf_table<<<nblocks(hw, 256), 256, 0, s>>>(hw, n, Sz, first, label, starttdc, *out);
  CRV_HIP(c, hipGetLastError());
  if (out->adc) k_copy_i16<<<nblocks(n * Sz, 256), 256, 0, s>>>(adc, out->adc, n * Sz);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

// ---- derived parameters ----
static Derived derive(const crv_params &p) {
  Derived d;
  double att = (p.w_short * fm_exp(-p.ref_distance / p.l_short)) + (p.w_long * fm_exp(-p.ref_distance / p.l_long));
  d.yscale = p.pe_per_mev / att;
  double pf = p.p1 * (1.0 - fm_exp(-p.overvoltage / p.p2));
  d.q_ct = (1.0 - sqrt(sqrt(1.0 - p.cross_talk))) / pf;
  d.nreach = (int32_t)ceil(p.single_pe_max_time / 12.5) + 2;
  return d;
}

}  // namespace crv

using namespace crv;

namespace {

int guard(crv_ctx *ctx) {
  if (!ctx) return CRV_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return set_error(&ctx->c, CRV_E_HIP, "hipSetDevice failed");
  return CRV_OK;
}

__global__ void k_gather_cond(int64_t nw, const int32_t *__restrict__ chan, const double *__restrict__ ped,
                              const double *__restrict__ area, const double *__restrict__ height,
                              const double *__restrict__ toff, double *__restrict__ o_ped, double *__restrict__ o_area,
                              double *__restrict__ o_height, double *__restrict__ o_toff) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  int ch = chan[w];
  o_ped[w] = ped[ch];
  o_area[w] = area[ch];
  o_height[w] = height[ch];
  o_toff[w] = toff[ch];
}

__global__ void k_bar_sipm(int64_t n, const int32_t *__restrict__ chan, int32_t *__restrict__ bar,
                           int32_t *__restrict__ sipm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bar[i] = chan[i] / 4;
  sipm[i] = chan[i] % 4;
}

}  // namespace

extern "C" {

int crv_ctx_create(int device, crv_ctx **out) {
  if (!out) return set_error(nullptr, CRV_E_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return set_error(nullptr, CRV_E_HIP, "no such HIP device");
  if (hipSetDevice(device) != hipSuccess) return set_error(nullptr, CRV_E_HIP, "hipSetDevice failed");
  crv_ctx *ctx = new crv_ctx();
  ctx->c.device = device;
  if (hipMalloc(&ctx->c.flags, sizeof(int) * F_NFLAGS) != hipSuccess) {
    delete ctx;
    return set_error(nullptr, CRV_E_HIP, "hipMalloc failed");
  }
  for (auto &e : ctx->c.ev) (void)hipEventCreate(&e);
  for (auto &e : ctx->c.kev) (void)hipEventCreate(&e);
  for (auto &e : ctx->c.fork) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
  for (auto &q : ctx->c.aux)
    if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
      (void)crv_ctx_destroy(ctx);
      return set_error(nullptr, CRV_E_HIP, "hipStreamCreate failed");
    }
  if (noalloc_setup(&ctx->c) != CRV_OK) {
    const std::string m = ctx->c.err;
    (void)crv_ctx_destroy(ctx);
    return set_error(nullptr, CRV_E_HIP, m);
  }
  *out = ctx;
  return CRV_OK;
}

int crv_ctx_destroy(crv_ctx *ctx) {
  if (!ctx) return CRV_OK;
  (void)hipSetDevice(ctx->c.device);
  (void)hipDeviceSynchronize();
  for (void *p : ctx->c.ws)
    if (p) (void)hipFree(p);
  for (void *p : ctx->c.owned) (void)hipFree(p);
  if (ctx->c.flags) (void)hipFree(ctx->c.flags);
  for (auto &e : ctx->c.ev) (void)hipEventDestroy(e);
  for (auto &e : ctx->c.kev) (void)hipEventDestroy(e);
  for (auto &e : ctx->c.fork) (void)hipEventDestroy(e);
  for (auto &e : ctx->c.pev) (void)hipEventDestroy(e);
  for (auto &q : ctx->c.aux)
    if (q) (void)hipStreamDestroy(q);
  delete ctx;
  return CRV_OK;
}

int crv_last_error(crv_ctx *ctx, char *buf, int64_t size) {
  if (!buf || size <= 0) return CRV_E_ARG;
  const std::string &m = ctx ? ctx->c.err : g_create_error;
  int64_t k = std::min<int64_t>((int64_t)m.size(), size - 1);
  std::memcpy(buf, m.data(), k);
  buf[k] = 0;
  return CRV_OK;
}

int crv_set_params(crv_ctx *ctx, const crv_params *p) {
  if (int rc = guard(ctx)) return rc;
  if (!p) return set_error(&ctx->c, CRV_E_ARG, "params is NULL");
  if (p->samples_zs <= 0 || p->n_pixels_x <= 0 || p->n_pixels_x + 1 > CRV_NPIX || p->n_pixels_y + 1 > CRV_NPIX)
    return set_error(&ctx->c, CRV_E_ARG, "bad SiPM or digi parameters");
  ctx->c.p = *p;
  ctx->c.d = derive(*p);
  ctx->c.have_params = true;
  return CRV_OK;
}

int crv_upload_conditions(crv_ctx *ctx, int32_t n, const int32_t *status, const int32_t *roc, const int32_t *feb,
                          const double *dev, const double *pedestal, const double *area, const double *height,
                          const double *toff) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (n <= 0 || !status || !roc || !feb || !dev || !pedestal || !area || !height || !toff)
    return set_error(c, CRV_E_ARG, "conditions: bad arguments");
  int rc;
  if ((rc = upload(c, status, n, &c->t.status)) || (rc = upload(c, roc, n, &c->t.roc)) ||
      (rc = upload(c, feb, n, &c->t.feb)) || (rc = upload(c, dev, n, &c->t.dev)) ||
      (rc = upload(c, pedestal, n, &c->t.pedestal)) || (rc = upload(c, area, n, &c->t.area)) ||
      (rc = upload(c, height, n, &c->t.height)) || (rc = upload(c, toff, n, &c->t.toff)))
    return rc;
  c->t.n_channels = n;
  return CRV_OK;
}

int crv_upload_template(crv_ctx *ctx, const double *tmpl, int32_t n) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (n <= 0 || !tmpl) return set_error(c, CRV_E_ARG, "template: bad arguments");
  int rc = upload(c, tmpl, n, &c->t.tmpl);
  if (rc) return rc;
  std::vector<float> tf(tmpl, tmpl + n);  // the FP32 batched chain's copy
  if ((rc = upload(c, tf.data(), n, &c->t.tmpl_f))) return rc;
  c->t.ntmpl = n;
  double m = tmpl[0];
  for (int i = 1; i < n; ++i) m = std::max(m, tmpl[i]);
  c->t.vmax = m;
  return CRV_OK;
}

int crv_upload_photon_map(crv_ctx *ctx, const double *pmap, int32_t nx, int32_t ny, const uint8_t *dead41) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (nx <= 0 || ny <= 0 || !pmap || !dead41) return set_error(c, CRV_E_ARG, "photon map: bad arguments");
  if (nx > CRV_NPIX - 1 || ny > CRV_NPIX - 1) return set_error(c, CRV_E_ARG, "photon map larger than 40 x 40 pixels");
  std::vector<double> cdf(nx * ny);
  double acc = 0.0;
  for (int i = 0; i < nx * ny; ++i) {
    acc = acc + pmap[i];
    cdf[i] = acc;
  }
  if (!(acc > 0)) return set_error(c, CRV_E_ARG, "photon map has no content");
  std::vector<uint32_t> dead(CRV_DEAD_WORDS, 0u);
  for (int i = 0; i < CRV_NPIX * CRV_NPIX; ++i)
    if (dead41[i]) dead[i >> 5] |= 1u << (i & 31);
  int rc;
  if ((rc = upload(c, cdf.data(), nx * ny, &c->t.cdf)) || (rc = upload(c, dead.data(), CRV_DEAD_WORDS, &c->t.dead)))
    return rc;
  int w_lo = CRV_DEAD_WORDS, w_hi = -1;
  for (int w = 0; w < CRV_DEAD_WORDS; ++w)
    if (dead[w]) {
      w_lo = std::min(w_lo, w);
      w_hi = w;
    }
  if (w_hi < 0) w_lo = 0;
  c->t.dead_w0 = w_hi - w_lo < 8 ? w_lo : -1;
  for (int k = 0; k < 8; ++k) c->t.dead_win[k] = c->t.dead_w0 >= 0 && w_lo + k < CRV_DEAD_WORDS ? dead[w_lo + k] : 0u;
  c->t.ncdf = nx * ny;
  c->t.map_ny = ny;
  c->t.map_uniform = 1;
  for (int i = 0; i < nx * ny; ++i)
    if (pmap[i] != 1.0) c->t.map_uniform = 0;
  c->t.cdf_total = acc;
  return CRV_OK;
}

int crv_upload_geometry(crv_ctx *ctx, int32_t nbars, const int32_t *sector, const int32_t *layer,
                        const int32_t *counter_in_type, const double *x, const double *y, const double *pos,
                        int32_t nsectors, const int32_t *sector_type, const int32_t *length_dir,
                        const double *pos_side0, const double *pos_side1, const uint8_t *sipms_side0,
                        const uint8_t *sipms_side1, const uint8_t *has_cuts, const double *cuts) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (nbars <= 0 || nsectors <= 0) return set_error(c, CRV_E_ARG, "geometry: bad sizes");
  int32_t ncit = 0;
  for (int i = 0; i < nbars; ++i) {
    if (sector[i] < 0 || sector[i] >= nsectors) return set_error(c, CRV_E_ARG, "geometry: bad sector index");
    if (layer[i] < 0 || layer[i] >= 4) return set_error(c, CRV_E_ARG, "geometry: layer must be 0..3");
    if (counter_in_type[i] < 0) return set_error(c, CRV_E_ARG, "geometry: negative counter_in_type");
    ncit = std::max(ncit, counter_in_type[i] + 1);
  }
  // sector-type limits (reco/coinc.py:112-122): over every sector of the type; types without cuts are dropped
  std::map<int, std::vector<int>> bytype;
  for (int s = 0; s < nsectors; ++s) bytype[sector_type[s]].push_back(s);
  std::vector<int32_t> tid;
  std::vector<double> tp0, tp1, thalf, tmaxdt;
  std::vector<uint8_t> ts0, ts1;
  for (auto &kv : bytype) {
    double mdt = -INFINITY;
    bool any = false;
    double p0 = INFINITY, p1 = -INFINITY;
    bool s0 = false, s1 = false;
    for (int s : kv.second) {
      p0 = std::min(p0, pos_side0[s]);
      p1 = std::max(p1, pos_side1[s]);
      s0 |= sipms_side0[s] != 0;
      s1 |= sipms_side1[s] != 0;
      if (has_cuts[s]) {
        any = true;
        mdt = std::max(mdt, cuts[s * 10 + 2]);
      }
    }
    if (!any) continue;
    tid.push_back(kv.first);
    tp0.push_back(p0);
    tp1.push_back(p1);
    thalf.push_back((p1 - p0) / 2);
    tmaxdt.push_back(mdt);
    ts0.push_back(s0);
    ts1.push_back(s1);
  }
  int rc;
  if ((rc = upload(c, sector, nbars, &c->t.sector)) || (rc = upload(c, layer, nbars, &c->t.layer)) ||
      (rc = upload(c, counter_in_type, nbars, &c->t.cit)) || (rc = upload(c, x, nbars, &c->t.bx)) ||
      (rc = upload(c, y, nbars, &c->t.by)) || (rc = upload(c, pos, (int64_t)nbars * 3, &c->t.bpos)) ||
      (rc = upload(c, sector_type, nsectors, &c->t.stype)) || (rc = upload(c, length_dir, nsectors, &c->t.ldir)) ||
      (rc = upload(c, pos_side0, nsectors, &c->t.ps0)) || (rc = upload(c, pos_side1, nsectors, &c->t.ps1)) ||
      (rc = upload(c, sipms_side0, nsectors, &c->t.sside0)) || (rc = upload(c, sipms_side1, nsectors, &c->t.sside1)) ||
      (rc = upload(c, has_cuts, nsectors, &c->t.hascuts)) || (rc = upload(c, cuts, (int64_t)nsectors * 10, &c->t.cuts)) ||
      (rc = upload(c, tid.data(), (int64_t)tid.size(), &c->t.type_id)) ||
      (rc = upload(c, tp0.data(), (int64_t)tp0.size(), &c->t.tp0)) ||
      (rc = upload(c, tp1.data(), (int64_t)tp1.size(), &c->t.tp1)) ||
      (rc = upload(c, thalf.data(), (int64_t)thalf.size(), &c->t.thalf)) ||
      (rc = upload(c, tmaxdt.data(), (int64_t)tmaxdt.size(), &c->t.tmaxdt)) ||
      (rc = upload(c, ts0.data(), (int64_t)ts0.size(), &c->t.tside0)) ||
      (rc = upload(c, ts1.data(), (int64_t)ts1.size(), &c->t.tside1)))
    return rc;
  std::vector<BarInfo> bi(nbars);
  for (int i = 0; i < nbars; ++i) {
    const int sc = sector[i];
    bi[i] = BarInfo{x[i], y[i], pos_side0[sc], pos_side1[sc], cuts[sc * 10 + 8], sc,
                    has_cuts[sc] ? sector_type[sc] : INT32_MIN, layer[i], counter_in_type[i], {0, 0}};
  }
  if ((rc = upload(c, bi.data(), nbars, &c->t.binfo))) return rc;
  c->t.nbars = nbars;
  c->t.nsectors = nsectors;
  c->t.ntypes = (int32_t)tid.size();
  c->t.ncit = ncit;
  return CRV_OK;
}

static int ready(Ctx *c, bool need_geom) {
  if (!c->have_params) return set_error(c, CRV_E_STATE, "parameters not set");
  if (!c->t.status) return set_error(c, CRV_E_STATE, "conditions not uploaded");
  if (!c->t.tmpl) return set_error(c, CRV_E_STATE, "template not uploaded");
  if (!c->t.cdf) return set_error(c, CRV_E_STATE, "photon map not uploaded");
  if (need_geom && !c->t.cuts) return set_error(c, CRV_E_STATE, "geometry not uploaded");
  return CRV_OK;
}

int crv_charges_run(crv_ctx *ctx, const crv_steps *steps, crv_charges *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !out) return set_error(c, CRV_E_ARG, "NULL argument");
  if (int rc = ready(c, false)) return rc;
  hipStream_t s = (hipStream_t)stream;
  StepGroups g;
  int rc = group_steps(c, steps, g, s);
  if (rc) return rc;
  return charges_stage(c, g, out, s);
}

int crv_digis_run(crv_ctx *ctx, const int64_t *label, const double *time, const double *charge, int64_t n,
                  crv_digis *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!out || (n && (!label || !time || !charge))) return set_error(c, CRV_E_ARG, "NULL argument");
  if (int rc = ready(c, false)) return rc;
  return digis_stage(c, label, time, charge, n, out, (hipStream_t)stream);
}

int crv_reco_waveforms_run(crv_ctx *ctx, const int64_t *label, const int64_t *starttdc, const int16_t *adc, int64_t n,
                           crv_waveforms *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!out || (n && (!label || !starttdc || !adc))) return set_error(c, CRV_E_ARG, "NULL argument");
  if (!c->have_params) return set_error(c, CRV_E_STATE, "parameters not set");
  return waveforms_stage(c, label, starttdc, adc, n, out, (hipStream_t)stream);
}

int crv_pulses_run(crv_ctx *ctx, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                   const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                   const double *height, const double *toff, crv_pulses *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!out || (n_wf && (!start || !nsamp || !off || !adc || !ped || !area || !height || !toff)))
    return set_error(c, CRV_E_ARG, "NULL argument");
  if (!c->have_params || !c->t.status) return set_error(c, CRV_E_STATE, "parameters or conditions missing");
  return pulses_stage(c, ev, chan, start, nsamp, off, adc, n_wf, ped, area, height, toff, out, (hipStream_t)stream);
}

int crv_pulses_run_noalloc(crv_ctx *ctx, const int64_t *ev, const int32_t *chan, const int64_t *start,
                           const int32_t *nsamp, const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped,
                           const double *area, const double *height, const double *toff, crv_pulses *out,
                           crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!out || (n_wf && (!start || !nsamp || !off || !adc || !ped || !area || !height || !toff)))
    return set_error(c, CRV_E_ARG, "NULL argument");
  if (n_wf < 0) return set_error(c, CRV_E_ARG, "pulses (no-alloc): negative waveform count");
  if (!c->have_params || !c->t.status) return set_error(c, CRV_E_STATE, "parameters or conditions missing");
  return pulses_stage_noalloc(c, ev, chan, start, nsamp, off, adc, n_wf, ped, area, height, toff, out,
                              (hipStream_t)stream);
}

int crv_clusters_run(crv_ctx *ctx, const int64_t *ev, const int32_t *bar, const int32_t *sipm, const double *time,
                     const double *pe, const int64_t *seq, int64_t n, crv_clusters *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!out || (n && (!ev || !bar || !sipm || !time || !pe || !seq))) return set_error(c, CRV_E_ARG, "NULL argument");
  if (!c->have_params || !c->t.status) return set_error(c, CRV_E_STATE, "parameters or conditions missing");
  return clusters_stage(c, ev, bar, sipm, time, pe, seq, n, out, (hipStream_t)stream);
}

int crv_chain_run(crv_ctx *ctx, const crv_steps *steps, crv_pulses *pulses, crv_clusters *clusters,
                  crv_timing *timing, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !pulses || !clusters) return set_error(c, CRV_E_ARG, "NULL argument");
  if (int rc = ready(c, true)) return rc;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t *ev = c->ev;
  c->kev_on = true;
  struct Off {
    Ctx *c;
    ~Off() { c->kev_on = false; }
  } off_{c};
  CRV_HIP(c, hipEventRecord(ev[0], s));
  StepGroups g;
  int rc = group_steps(c, steps, g, s);
  if (rc) return rc;
  crv_charges q{};
  q.capacity = -1;
  if ((rc = charges_stage(c, g, &q, s))) return rc;
  CRV_HIP(c, hipEventRecord(ev[1], s));
  crv_digis d{};
  d.capacity = -1;
  if ((rc = digis_stage(c, q.label, q.time, q.charge, q.n, &d, s))) return rc;
  CRV_HIP(c, hipEventRecord(ev[2], s));
  crv_waveforms w{};
  w.capacity = -1;
  if ((rc = waveforms_stage(c, d.label, d.starttdc, d.adc, d.n, &w, s))) return rc;
  double *wped = (double *)ws(c, 310, w.n * 8 + 8), *warea = (double *)ws(c, 311, w.n * 8 + 8);
  double *wheight = (double *)ws(c, 312, w.n * 8 + 8), *wtoff = (double *)ws(c, 313, w.n * 8 + 8);
  if (!wped || !warea || !wheight || !wtoff) return ws_fail(c);
  if (w.n) {
    k_gather_cond<<<nblocks(w.n, 256), 256, 0, s>>>(w.n, w.chan, c->t.pedestal, c->t.area, c->t.height, c->t.toff,
                                                    wped, warea, wheight, wtoff);
    CRV_HIP(c, hipGetLastError());
  }
  CRV_HIP(c, hipEventRecord(ev[3], s));
  crv_pulses P{};
  P.capacity = -1;
  if ((rc = pulses_stage(c, w.ev, w.chan, w.start, w.nsamp, w.off, d.adc, w.n, wped, warea, wheight, wtoff, &P, s)))
    return rc;
  CRV_HIP(c, hipEventRecord(ev[4], s));
  int32_t *pbar = (int32_t *)ws(c, 316, P.n * 4 + 8), *psipm = (int32_t *)ws(c, 317, P.n * 4 + 8);
  if (!pbar || !psipm) return ws_fail(c);
  if (P.n) {
    k_bar_sipm<<<nblocks(P.n, 256), 256, 0, s>>>(P.n, P.chan, pbar, psipm);
    CRV_HIP(c, hipGetLastError());
  }
  crv_clusters C{};
  C.capacity = -1;
  if ((rc = clusters_stage(c, P.ev, pbar, psipm, P.time, P.pe, P.seq, P.n, &C, s))) return rc;
  CRV_HIP(c, hipEventRecord(ev[5], s));
  CRV_HIP(c, hipEventSynchronize(ev[5]));
  if (timing) {
    float ms[5];
    for (int i = 0; i < 5; ++i) (void)hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
    timing->charges_ms = ms[0];
    timing->digis_ms = ms[1];
    timing->waveforms_ms = ms[2];
    timing->pulses_ms = ms[3];
    timing->clusters_ms = ms[4];
    float tot = 0;
    (void)hipEventElapsedTime(&tot, ev[0], ev[5]);
    timing->total_ms = tot;
    float km[4] = {0, 0, 0, 0};
    bool ran[4] = {g.T > 0, q.n > 0, w.n > 0, P.n > 0};
    for (int k = 0; k < 4; ++k)
      if (ran[k]) (void)hipEventElapsedTime(&km[k], c->kev[2 * k], c->kev[2 * k + 1]);
    timing->k_sipm_ms = km[0];
    timing->k_digi_ms = km[1];
    timing->k_pulse_ms = km[2];
    timing->k_coinc_ms = km[3];
    timing->n_steps = steps->n;
    timing->n_hit_counters = g.T;
    timing->n_photons = q.n_photons;
    timing->n_charges = q.n;
    timing->n_digis = d.n;
    timing->n_records = d.n_records;
    timing->n_waveforms = w.n;
    timing->n_pulses = P.n;
    timing->n_clusters = C.n;
    timing->n_members = C.n_mem;
  }
  c->last_digis = d;
  c->last_pulses = P;
  c->last_clusters = C;
  return crv_chain_fetch(ctx, pulses, clusters, stream);
}

int crv_chain_run_batched(crv_ctx *ctx, const crv_steps *steps, crv_batch_opts *opts, crv_pulses *pulses,
                          crv_clusters *clusters, crv_timing *timing, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !opts || !pulses || !clusters) return set_error(c, CRV_E_ARG, "NULL argument");
  if (int rc = ready(c, true)) return rc;
  c->kev_on = true;
  struct Off {
    Ctx *c;
    ~Off() { c->kev_on = false; }
  } off_{c};
  c->last_digis = crv_digis{};
  int rc = batched_stage(c, steps, opts, timing, (hipStream_t)stream);
  if (rc && rc != CRV_E_CAPACITY) return rc;
  int rf = crv_chain_fetch(ctx, pulses, clusters, stream);
  return rf ? rf : rc;
}

int crv_chain_records(crv_ctx *ctx, const uint16_t **adc, const int64_t **label, int64_t *n) {
  if (int rc = guard(ctx)) return rc;
  if (!adc || !label || !n) return set_error(&ctx->c, CRV_E_ARG, "NULL argument");
  *adc = ctx->c.last_digis.record_adc;
  *label = ctx->c.last_digis.record_label;
  *n = ctx->c.last_digis.n_records;
  return CRV_OK;
}

int crv_chain_fetch(crv_ctx *ctx, crv_pulses *pulses, crv_clusters *clusters, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  hipStream_t s = (hipStream_t)stream;
  const crv_pulses &P = c->last_pulses;
  const crv_clusters &C = c->last_clusters;
  pulses->n = P.n;
  clusters->n = C.n;
  clusters->n_mem = C.n_mem;
  if (P.n > pulses->capacity || C.n > clusters->capacity || C.n_mem > clusters->mem_capacity)
    return set_error(c, CRV_E_CAPACITY, "chain: pulse or cluster capacity too small (sizes in n; call crv_chain_fetch)");
  auto cp = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
    if (!dst || !bytes) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
  };
  const int64_t np = P.n, nc = C.n;
  hipError_t e = hipSuccess;
  void *pd[] = {pulses->wf, pulses->seq, pulses->ev};
  const void *ps[] = {P.wf, P.seq, P.ev};
  for (int i = 0; i < 3; ++i) e = e == hipSuccess ? cp(pd[i], ps[i], np * 8) : e;
  e = e == hipSuccess ? cp(pulses->chan, P.chan, np * 4) : e;
  double *pdd[] = {pulses->pe, pulses->peph, pulses->time, pulses->height, pulses->beta, pulses->chi2, pulses->letime,
                   pulses->ped, pulses->fit_a, pulses->fit_mu, pulses->fit_b};
  const double *psd[] = {P.pe, P.peph, P.time, P.height, P.beta, P.chi2, P.letime, P.ped, P.fit_a, P.fit_mu, P.fit_b};
  for (int i = 0; i < 11; ++i) e = e == hipSuccess ? cp(pdd[i], psd[i], np * 8) : e;
  e = e == hipSuccess ? cp(pulses->failed, P.failed, np) : e;
  e = e == hipSuccess ? cp(pulses->zerondf, P.zerondf, np) : e;
  e = e == hipSuccess ? cp(pulses->pedfromdb, P.pedfromdb, np) : e;
  e = e == hipSuccess ? cp(clusters->cl_event, C.cl_event, nc * 8) : e;
  e = e == hipSuccess ? cp(clusters->cl_type, C.cl_type, nc * 4) : e;
  e = e == hipSuccess ? cp(clusters->cl_poscalc, C.cl_poscalc, nc * 4) : e;
  e = e == hipSuccess ? cp(clusters->cl_sidehits, C.cl_sidehits, nc * 8) : e;
  double *cdd[] = {clusters->cl_pe, clusters->cl_start, clusters->cl_end, clusters->cl_slope, clusters->cl_avgtime};
  const double *csd[] = {C.cl_pe, C.cl_start, C.cl_end, C.cl_slope, C.cl_avgtime};
  for (int i = 0; i < 5; ++i) e = e == hipSuccess ? cp(cdd[i], csd[i], nc * 8) : e;
  e = e == hipSuccess ? cp(clusters->cl_pos, C.cl_pos, nc * 24) : e;
  e = e == hipSuccess ? cp(clusters->cl_sidepe, C.cl_sidepe, nc * 16) : e;
  e = e == hipSuccess ? cp(clusters->cl_sidetime, C.cl_sidetime, nc * 16) : e;
  e = e == hipSuccess ? cp(clusters->cl_mem_off, C.cl_mem_off, (nc + 1) * 8) : e;
  e = e == hipSuccess ? cp(clusters->mem_idx, C.mem_idx, C.n_mem * 8) : e;
  if (e != hipSuccess) return set_error(c, CRV_E_HIP, std::string("chain fetch: ") + hipGetErrorString(e));
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

int crv_ctx_workspace(crv_ctx *ctx, uint64_t *bytes) {
  if (int rc = guard(ctx)) return rc;
  if (!bytes) return set_error(&ctx->c, CRV_E_ARG, "NULL argument");
  uint64_t b = 0;
  for (size_t i = 0; i < ctx->c.ws.size(); ++i)
    if (ctx->c.ws[i]) b += ctx->c.ws_size[i];
  *bytes = b;
  return CRV_OK;
}

int crv_ctx_trim(crv_ctx *ctx, uint64_t *freed) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  CRV_HIP(c, hipDeviceSynchronize());  // no kernel of this context still reads a slot
  uint64_t b = 0;
  for (size_t i = 0; i < c->ws.size(); ++i)
    if (c->ws[i]) {
      (void)hipFree(c->ws[i]);
      b += c->ws_size[i];
      c->ws[i] = nullptr;
      c->ws_size[i] = 0;
    }
  c->ws_spare.clear();
  c->last_pulses = crv_pulses{};
  c->last_clusters = crv_clusters{};
  c->last_digis = crv_digis{};
  c->b_rec_cap = c->b_run_cap = 0;
  if (freed) *freed = b;
  return CRV_OK;
}

}  // extern "C"
