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

// a captured FP32 batched chain (crv_graph_*): the plan's buffers, the graph and its launch statistics
struct crv_graph {
  crv_ctx *ctx = nullptr;
  crv::Plan pl;
  crv::Arena fixed;   // inputs, counts and outputs of the plan
  crv::Arena run;     // everything the captured sequence uses
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cs = nullptr;  // the capture stream
  hipEvent_t e0 = nullptr, e1 = nullptr;
  crv_params params{};        // the parameters captured into the kernels
  crv_batch_opts opts{};      // queue options of the plan (for the uncaptured runs)
  int64_t launches = 0, fallbacks = 0;
  int last_fell_back = 0;
  int64_t sizes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t n_host = 0;  // the step count a launch copies to the device (kept alive until the copy has run)
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

void *Arena::get(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  if (replay) {  // the same request order as the eager run: hand out its buffers again
    if (next >= blocks.size() || blocks[next].second < bytes) return nullptr;
    return blocks[next++].first;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    failed = bytes;
    return nullptr;
  }
  blocks.emplace_back(p, bytes);
  return p;
}

std::string Arena::note() const {
  size_t held = 0, fr = 0, total = 0;
  for (const auto &b : blocks) held += b.second;
  (void)hipMemGetInfo(&fr, &total);
  return " (request " + std::to_string(failed >> 20) + " MiB, the graph holds " + std::to_string(held >> 20) +
         " MiB, device free " + std::to_string(fr >> 20) + " of " + std::to_string(total >> 20) + " MiB)";
}

void Arena::release() {
  for (auto &b : blocks) (void)hipFree(b.first);
  blocks.clear();
  next = 0;
  replay = false;
}

namespace {
__global__ void k_first_zero(int64_t *__restrict__ out) { *out = 0; }
}  // namespace

int scan_fixed(Ctx *c, Arena &ar, const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
  k_first_zero<<<1, 1, 0, s>>>(out);
  CRV_HIP(c, hipGetLastError());
  if (n == 0) return CRV_OK;
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out + 1, (int)n, s));
  void *tmp = ar.get(tb);
  if (!tmp) return set_error(c, CRV_E_HIP, "graph: scan buffer allocation failed" + ar.note());
  CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out + 1, (int)n, s));
  return CRV_OK;
}

int widen_scan_fixed(Ctx *c, Arena &ar, const int32_t *in, int64_t *out, int64_t n, hipStream_t s) {
  int64_t *w = (int64_t *)ar.get((n + 1) * 8);
  if (!w) return set_error(c, CRV_E_HIP, "graph: scan buffer allocation failed" + ar.note());
  if (n) k_i32_to_i64<<<nblocks(n, 256), 256, 0, s>>>(in, w, n);
  CRV_HIP(c, hipGetLastError());
  return scan_fixed(c, ar, w, out, n, s);
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

// copy a pulse and cluster table (device) into the caller's buffers; CRV_E_CAPACITY with the sizes when they are short
int copy_tables(Ctx *c, const crv_pulses &P, const crv_clusters &C, crv_pulses *pulses, crv_clusters *clusters,
                hipStream_t s) {
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
  lookup_free(&ctx->c);
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

int crv_photons_run(crv_ctx *ctx, const crv_steps *steps, const double *shift, double period, crv_photons *out,
                    crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !out) return set_error(c, CRV_E_ARG, "NULL argument");
  if (int rc = ready(c, false)) return rc;
  hipStream_t s = (hipStream_t)stream;
  StepGroups g;
  int rc = group_steps(c, steps, g, s);
  if (rc) return rc;
  return photons_stage(c, g, shift, period, out, s);
}

int crv_lookup_upload(crv_ctx *ctx, const crv_lookup_table *tables, int32_t n_tables, const crv_lookup_bars *bars) {
  if (int rc = guard(ctx)) return rc;
  return lookup_upload(&ctx->c, tables, n_tables, bars);
}

int crv_lookup_photons_run(crv_ctx *ctx, const crv_lookup_steps *steps, crv_photons *out, crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !out) return set_error(c, CRV_E_ARG, "NULL argument");
  if (!c->have_params) return set_error(c, CRV_E_STATE, "parameters not set");
  if (!c->t.status) return set_error(c, CRV_E_STATE, "conditions not uploaded");
  return lookup_photons_stage(c, steps, out, (hipStream_t)stream);
}

int crv_sipm_run(crv_ctx *ctx, const crv_photon_list *photons, const crv_sipm_noise *noise, crv_charges *out,
                 crv_stream stream) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!photons || !noise || !out) return set_error(c, CRV_E_ARG, "NULL argument");
  if (!c->have_params) return set_error(c, CRV_E_STATE, "parameters not set");
  if (!c->t.cdf) return set_error(c, CRV_E_STATE, "photon map not uploaded");
  return sipm_list_stage(c, photons, noise, out, (hipStream_t)stream);
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
  return copy_tables(c, c->last_pulses, c->last_clusters, pulses, clusters, (hipStream_t)stream);
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
  c->b_rec_cap = c->b_run_cap = c->b_desc_cap = 0;
  c->b_pcap = 0;
  if (freed) *freed = b;
  return CRV_OK;
}


// ---- captured batched chain ----

int crv_graph_create(crv_ctx *ctx, const crv_steps *sample, int64_t max_steps, const crv_batch_opts *opts,
                     double margin, crv_graph **out) {
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!sample || !opts || !out) return set_error(c, CRV_E_ARG, "NULL argument");
  *out = nullptr;
  if (int rc = ready(c, true)) return rc;
  if (max_steps < sample->n || max_steps <= 0 || max_steps > INT32_MAX / 4)
    return set_error(c, CRV_E_ARG, "graph: max_steps must hold the sample and stay below 2^29");
  if (!(margin >= 0.0) || margin > 10.0) return set_error(c, CRV_E_ARG, "graph: margin must be in [0, 10]");
  if (opts->tile <= 0 || opts->tile % 64 || opts->sort_tasks < 0 || opts->sort_tasks > 3)
    return set_error(c, CRV_E_ARG, "graph: opts need a tile (multiple of 64) and a sort order 0-3");
  // sizing: the uncaptured chain on the sample gives the workload's table sizes per step
  crv_batch_opts so = *opts;
  so.log_budget = 0;
  so.pipeline = 0;
  crv_timing tm{};
  c->kev_on = false;
  int rc = batched_stage(c, sample, &so, &tm, nullptr);
  if (rc) return rc;
  const double scale = (double)max_steps / (double)std::max<int64_t>(sample->n, 1) * (1.0 + margin);
  auto cap = [&](int64_t v) { return (int64_t)std::ceil((double)v * scale) + 1024; };
  crv_graph *G = new crv_graph();
  G->ctx = ctx;
  G->opts = so;
  Plan &pl = G->pl;
  pl.n_cap = max_steps;
  pl.log_cap = cap(c->b_last_log);
  pl.rec_cap = cap(tm.n_records);
  pl.run_cap = cap(tm.n_waveforms);
  pl.np_cap = cap(tm.n_pulses);
  pl.desc_cap = cap(c->b_last_desc);
  pl.pcap = std::max(so.pulse_slots_used, 4);
  pl.log_factor = c->b_log_factor;
  pl.tile = so.tile;
  pl.sort_tasks = so.sort_tasks;
  pl.grid_s = pl.grid_r = so.grid;
  // the graph takes every buffer from its own arenas: the sizing run's workspace goes back to the device first, so
  // the device never holds two copies of a call's buffers (an uncaptured fallback allocates it again)
  if ((rc = crv_ctx_trim(ctx, nullptr))) return rc;
  auto fail = [&](int code, const std::string &m) {
    (void)crv_graph_destroy(G);
    return set_error(c, code, m);
  };
  Arena &F = G->fixed;
  const int64_t n = pl.n_cap, np = pl.np_cap;
  pl.in_event = (uint32_t *)F.get(n * 4);
  pl.in_counter = (uint32_t *)F.get(n * 4);
  pl.in_edep = (double *)F.get(n * 8);
  pl.in_z = (double *)F.get(n * 8);
  pl.in_t = (double *)F.get(n * 8);
  pl.n_dev = (int64_t *)F.get(8);
  pl.T_dev = (int32_t *)F.get(8);
  pl.skip = (int *)F.get(16);
  pl.sizes = (int64_t *)F.get(64);
  crv_pulses &P = pl.P;
  int64_t **ip[] = {&P.wf, &P.seq, &P.ev};
  for (auto q : ip) *q = (int64_t *)F.get(np * 8 + 8);
  P.chan = (int32_t *)F.get(np * 4 + 8);
  double **dp[] = {&P.pe, &P.peph, &P.time, &P.height, &P.beta, &P.chi2, &P.letime, &P.ped, &P.fit_a, &P.fit_mu,
                   &P.fit_b};
  for (auto q : dp) *q = (double *)F.get(np * 8 + 8);
  uint8_t **bp[] = {&P.failed, &P.zerondf, &P.pedfromdb};
  for (auto q : bp) *q = (uint8_t *)F.get(np + 8);
  crv_clusters &C = pl.C;
  C.cl_event = (int64_t *)F.get(np * 8 + 8);
  C.cl_type = (int32_t *)F.get(np * 4 + 8);
  C.cl_poscalc = (int32_t *)F.get(np * 4 + 8);
  C.cl_sidehits = (int32_t *)F.get(np * 8 + 8);
  double **cp[] = {&C.cl_pe, &C.cl_start, &C.cl_end, &C.cl_slope, &C.cl_avgtime};
  for (auto q : cp) *q = (double *)F.get(np * 8 + 8);
  C.cl_pos = (double *)F.get(np * 24 + 8);
  C.cl_sidepe = (double *)F.get(np * 16 + 8);
  C.cl_sidetime = (double *)F.get(np * 16 + 8);
  C.cl_mem_off = (int64_t *)F.get((np + 2) * 8);
  C.mem_idx = (int64_t *)F.get(np * 8 + 8);
  bool ok = pl.in_event && pl.in_counter && pl.in_edep && pl.in_z && pl.in_t && pl.n_dev && pl.T_dev && pl.skip &&
            pl.sizes && P.chan && C.cl_event && C.cl_type && C.cl_poscalc && C.cl_sidehits && C.cl_pos && C.cl_sidepe &&
            C.cl_sidetime && C.cl_mem_off && C.mem_idx;
  for (auto q : ip) ok = ok && *q;
  for (auto q : dp) ok = ok && *q;
  for (auto q : bp) ok = ok && *q;
  for (auto q : cp) ok = ok && *q;
  if (!ok) return fail(CRV_E_HIP, "graph: plan buffer allocation failed" + F.note());
  if (hipStreamCreateWithFlags(&G->cs, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&G->e0) != hipSuccess || hipEventCreate(&G->e1) != hipSuccess)
    return fail(CRV_E_HIP, "graph: stream or event creation failed");
  for (auto &e : pl.kev)
    if (hipEventCreate(&e) != hipSuccess) return fail(CRV_E_HIP, "graph: event creation failed");
  // the eager run on the sample allocates the sequence's buffers and checks it against the uncaptured chain
  auto load = [&](const crv_steps *st, hipStream_t s) -> hipError_t {
    const int64_t m = st->n;
    G->n_host = m;
    hipError_t e = hipMemcpyAsync(pl.n_dev, &G->n_host, 8, hipMemcpyHostToDevice, s);
    if (m) {
      if (e == hipSuccess && st->event != pl.in_event)
        e = hipMemcpyAsync(pl.in_event, st->event, m * 4, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess && st->counter != pl.in_counter)
        e = hipMemcpyAsync(pl.in_counter, st->counter, m * 4, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess && st->edep != pl.in_edep)
        e = hipMemcpyAsync(pl.in_edep, st->edep, m * 8, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess && st->z != pl.in_z) e = hipMemcpyAsync(pl.in_z, st->z, m * 8, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess && st->t != pl.in_t) e = hipMemcpyAsync(pl.in_t, st->t, m * 8, hipMemcpyDeviceToDevice, s);
    }
    return e;
  };
  if (load(sample, G->cs) != hipSuccess) return fail(CRV_E_HIP, "graph: input copy failed");
  if ((rc = batched_fixed(c, pl, G->run, G->cs))) {
    (void)hipStreamSynchronize(G->cs);
    const std::string m = c->err;
    return fail(rc, m);
  }
  int skip[2] = {0, 0};
  if (hipMemcpyAsync(skip, pl.skip, 8, hipMemcpyDeviceToHost, G->cs) != hipSuccess ||
      hipMemcpyAsync(G->sizes, pl.sizes, 64, hipMemcpyDeviceToHost, G->cs) != hipSuccess ||
      hipStreamSynchronize(G->cs) != hipSuccess)
    return fail(CRV_E_HIP, "graph: eager run failed");
  if (skip[0]) return fail(CRV_E_STATE, "graph: the sample exceeds the plan's capacities (reason " +
                                          std::to_string(skip[1]) + ")");
  if (G->sizes[0] != tm.n_pulses || G->sizes[1] != tm.n_clusters || G->sizes[2] != tm.n_members)
    return fail(CRV_E_STATE, "graph: the fixed-capacity sequence disagrees with the batched chain on the sample");
  // capture the same sequence: it receives the eager run's buffers again, in the same order
  G->run.replay = true;
  G->run.next = 0;
  if (hipStreamBeginCapture(G->cs, hipStreamCaptureModeRelaxed) != hipSuccess)
    return fail(CRV_E_HIP, "graph: hipStreamBeginCapture failed");
  rc = batched_fixed(c, pl, G->run, G->cs);
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(G->cs, &gr);
  G->graph = gr;
  if (rc) {
    const std::string m = c->err;
    return fail(rc, "graph: capture: " + m);
  }
  if (ec != hipSuccess || !gr) return fail(CRV_E_HIP, std::string("graph: capture failed: ") + hipGetErrorString(ec));
  if (G->run.next != G->run.blocks.size()) return fail(CRV_E_STATE, "graph: the captured sequence's buffers differ");
  if (hipGraphInstantiate(&G->exec, gr, nullptr, nullptr, 0) != hipSuccess)
    return fail(CRV_E_HIP, "graph: hipGraphInstantiate failed");
  G->params = c->p;
  *out = G;
  return CRV_OK;
}

int crv_graph_launch(crv_graph *G, const crv_steps *steps, crv_pulses *pulses, crv_clusters *clusters,
                     crv_timing *timing, crv_stream stream) {
  if (!G) return CRV_E_ARG;
  crv_ctx *ctx = G->ctx;
  if (int rc = guard(ctx)) return rc;
  Ctx *c = &ctx->c;
  if (!steps || !pulses || !clusters) return set_error(c, CRV_E_ARG, "NULL argument");
  if (std::memcmp(&G->params, &c->p, sizeof(crv_params)) != 0)
    return set_error(c, CRV_E_STATE, "graph: parameters changed since the capture (create the graph again)");
  hipStream_t s = (hipStream_t)stream;
  ++G->launches;
  G->last_fell_back = 0;
  auto fallback = [&]() {
    ++G->fallbacks;
    G->last_fell_back = 1;
    crv_batch_opts o = G->opts;
    return crv_chain_run_batched(ctx, steps, &o, pulses, clusters, timing, stream);
  };
  if (steps->n > G->pl.n_cap) return fallback();
  Plan &pl = G->pl;
  const int64_t m = steps->n;
  G->n_host = m;
  CRV_HIP(c, hipMemcpyAsync(pl.n_dev, &G->n_host, 8, hipMemcpyHostToDevice, s));
  if (m) {
    if (steps->event != pl.in_event)
      CRV_HIP(c, hipMemcpyAsync(pl.in_event, steps->event, m * 4, hipMemcpyDeviceToDevice, s));
    if (steps->counter != pl.in_counter)
      CRV_HIP(c, hipMemcpyAsync(pl.in_counter, steps->counter, m * 4, hipMemcpyDeviceToDevice, s));
    if (steps->edep != pl.in_edep) CRV_HIP(c, hipMemcpyAsync(pl.in_edep, steps->edep, m * 8, hipMemcpyDeviceToDevice, s));
    if (steps->z != pl.in_z) CRV_HIP(c, hipMemcpyAsync(pl.in_z, steps->z, m * 8, hipMemcpyDeviceToDevice, s));
    if (steps->t != pl.in_t) CRV_HIP(c, hipMemcpyAsync(pl.in_t, steps->t, m * 8, hipMemcpyDeviceToDevice, s));
  }
  CRV_HIP(c, hipEventRecord(G->e0, s));
  CRV_HIP(c, hipGraphLaunch(G->exec, s));
  CRV_HIP(c, hipEventRecord(G->e1, s));
  int skip[2] = {0, 0}, hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(skip, pl.skip, 8, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipMemcpyAsync(G->sizes, pl.sizes, 64, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_UNSORTED]) return set_error(c, CRV_E_ARG, "steps must be sorted by event");
  if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "batched chain: channel beyond the uploaded conditions");
  if (hf[F_COUNTER])
    return set_error(c, CRV_E_OVERFLOW, "batched chain: counter id >= 8192 or more than 65535 photoelectrons from one "
                                        "step");
  if (hf[F_CASCADE])
    return set_error(c, CRV_E_OVERFLOW,
                     "batched chain: cross-talk cascade wider than 16 pixels or deeper than 31 generations");
  if (skip[0]) return fallback();  // a capacity was exceeded: this call runs uncaptured
  if (timing) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, G->e0, G->e1);
    *timing = crv_timing{};
    timing->total_ms = ms;
    float km[3] = {0.f, 0.f, 0.f};  // the finder's kernels run on three streams: no single span is measured
    for (int k = 0; k < 3; ++k) (void)hipEventElapsedTime(&km[k], pl.kev[2 * k], pl.kev[2 * k + 1]);
    timing->k_sipm_ms = km[0];
    timing->k_digi_ms = km[1];
    timing->k_pulse_ms = km[2];
    timing->n_steps = m;
    timing->n_photons = G->sizes[3];
    timing->n_charges = G->sizes[4];
    timing->n_records = G->sizes[5];
    timing->n_waveforms = G->sizes[6];
    timing->n_pulses = G->sizes[0];
    timing->n_clusters = G->sizes[1];
    timing->n_members = G->sizes[2];
  }
  return crv_graph_fetch(G, pulses, clusters, stream);
}

int crv_graph_fetch(crv_graph *G, crv_pulses *pulses, crv_clusters *clusters, crv_stream stream) {
  if (!G) return CRV_E_ARG;
  if (int rc = guard(G->ctx)) return rc;
  Ctx *c = &G->ctx->c;
  if (!pulses || !clusters) return set_error(c, CRV_E_ARG, "NULL argument");
  if (G->last_fell_back)
    return copy_tables(c, c->last_pulses, c->last_clusters, pulses, clusters, (hipStream_t)stream);
  crv_pulses P = G->pl.P;
  crv_clusters C = G->pl.C;
  P.n = G->sizes[0];
  C.n = G->sizes[1];
  C.n_mem = G->sizes[2];
  return copy_tables(c, P, C, pulses, clusters, (hipStream_t)stream);
}

int crv_graph_stats(crv_graph *G, int64_t *launches, int64_t *fallbacks, int64_t *capacities) {
  if (!G) return CRV_E_ARG;
  if (launches) *launches = G->launches;
  if (fallbacks) *fallbacks = G->fallbacks;
  if (capacities) {
    const Plan &pl = G->pl;
    const int64_t v[7] = {pl.n_cap, pl.log_cap, pl.rec_cap, pl.run_cap, pl.np_cap, pl.pcap, G->last_fell_back};
    for (int i = 0; i < 7; ++i) capacities[i] = v[i];
  }
  return CRV_OK;
}

int crv_graph_destroy(crv_graph *G) {
  if (!G) return CRV_OK;
  if (G->ctx) (void)hipSetDevice(G->ctx->c.device);
  if (G->cs) (void)hipStreamSynchronize(G->cs);
  (void)hipDeviceSynchronize();
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  if (G->e0) (void)hipEventDestroy(G->e0);
  if (G->e1) (void)hipEventDestroy(G->e1);
  for (auto &e : G->pl.kev)
    if (e) (void)hipEventDestroy(e);
  if (G->cs) (void)hipStreamDestroy(G->cs);
  G->run.release();
  G->fixed.release();
  delete G;
  return CRV_OK;
}

}  // extern "C"
