// Response stage: CrvSteps -> photoelectrons at the SiPMs -> SiPM charges (FP64, one launch per stage).
//
// Follows the reference's response/photons.py:102-224 (one count draw per (step, SiPM), arrival = step time +
// propagation + emission delay) with the closed-form optics, and response/sipm.py:102-198 (pixel from the photon
// map, recovery v = Vov (1 - e^{-dt/tau}), firing p1 (1 - e^{-v/p2}), cross-talk cascade breadth-first with
// multiplicity) and the status mask of chain.py:199-204.
//
// Layout: steps are grouped into hit counters (event, counter) by a radix sort of (event << 13 | counter); one lane
// simulates one SiPM (hit counter x 4), merging the time-ordered photon streams of the counter's steps (each
// stream's delays are generated in ascending order as exponential order statistics), so no photon sort is needed.
// The lane keeps an "ever fired" bitmap of its 41 x 41 pixels in LDS and looks a fired pixel's last discharge up
// in its own charge records (written in time order), newest first.
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "crv_internal.h"
#include "crv_sipm_dev.h"

namespace crv {

namespace {


__global__ void k_step_keys(const uint32_t *__restrict__ ev, const uint32_t *__restrict__ counter, int64_t n,
                            uint64_t *__restrict__ key, int32_t *__restrict__ idx, int32_t *__restrict__ evfirst,
                            int *__restrict__ flags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t e = ev[i];
  int64_t lo = 0, hi = i;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (ev[mid] < e) lo = mid + 1; else hi = mid;
  }
  evfirst[i] = (int32_t)lo;
  if (i > 0 && ev[i - 1] > e) atomicAdd(&flags[F_UNSORTED], 1);
  uint32_t c = counter[i];
  if (c >= 8192u) atomicAdd(&flags[F_COUNTER], 1);
  key[i] = ((uint64_t)e << 13) | (uint64_t)(c & 8191u);
  idx[i] = (int32_t)i;
}

struct Steps {
  const uint32_t *ev, *counter;
  const double *edep, *z, *t;
};

__device__ __forceinline__ double attenuation(const crv_params &p, double d) {
  return (p.w_short * fm_exp(-d / p.l_short)) + (p.w_long * fm_exp(-d / p.l_long));
}

__device__ __forceinline__ double dist_to_end(const crv_params &p, double z, int s) {
  return (s & 1) == 0 ? z + p.half_length : p.half_length - z;
}

// photoelectron count per (step, SiPM) and per SiPM
__global__ void k_photon_count(crv_params p, Tables t, Derived dv, Steps st, int64_t T,
                               const uint64_t *__restrict__ task_key, const int32_t *__restrict__ task_off,
                               const int32_t *__restrict__ perm, const int32_t *__restrict__ evfirst,
                               int32_t *__restrict__ nph, int64_t *__restrict__ sipm_nph, int *__restrict__ flags) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= T * 4) return;
  int64_t tk = g >> 2;
  int s = (int)(g & 3);
  uint64_t key = task_key[tk];
  uint32_t ev = (uint32_t)(key >> 13);
  int ch = (int)(key & 8191u) * 4 + s;
  int a = task_off[tk], b = task_off[tk + 1];
  bool live = ch < t.n_channels && (t.status[ch] & 0b101) == 0;
  if (ch >= t.n_channels) atomicAdd(&flags[F_BADCHAN], 1);
  Draw dr{(uint32_t)p.seed, p.evkey_hi};
  int64_t total = 0;
  for (int k = a; k < b; ++k) {
    int32_t n = 0;
    if (live) {
      int32_t i = perm[k];
      uint32_t j = (uint32_t)(i - evfirst[i]);
      double d = dist_to_end(p, st.z[i], s);
      double mu = ((st.edep[i] * dv.yscale) * attenuation(p, d)) * t.dev[ch];
      bool capped = false;
      long long nn = poisson(mu, dr, KIND_COUNT, ev, (uint32_t)ch, j, &capped);
      if (capped) atomicAdd(&flags[F_CAPPED], 1);
      if (nn > 65535) {  // photon ranks carry 16 bits in the draw layout
        atomicAdd(&flags[F_OVERFLOW], 1);
        nn = 65535;
      }
      n = (int32_t)nn;
    }
    nph[(int64_t)k * 4 + s] = n;
    total += n;
  }
  sipm_nph[g] = total;
}

__global__ void k_charge_caps(const int64_t *__restrict__ sipm_nph, int64_t n, int64_t *__restrict__ cap, int factor) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  int64_t v = sipm_nph[g];
  cap[g] = v + (v * factor) / 4 + 8;
}

struct Gen {  // one photon stream of a (step, SiPM): ordered exponential delays
  double *S, *head, *tb;
  int32_t *r;
};



__global__ __launch_bounds__(64) void k_sipm(crv_params p, Tables t, Derived dv, Steps st, int64_t T,
                                             const uint64_t *__restrict__ task_key,
                                             const int32_t *__restrict__ task_off, const int32_t *__restrict__ perm,
                                             const int32_t *__restrict__ evfirst, const int32_t *__restrict__ nph,
                                             const int64_t *__restrict__ coff, Gen gen, double *__restrict__ c_time,
                                             double *__restrict__ c_v, int32_t *__restrict__ c_row,
                                             int16_t *__restrict__ c_pix, int32_t *__restrict__ c_count,
                                             int *__restrict__ flags) {
  __shared__ SipmLds L;
  const int lane = threadIdx.x;
  int64_t g = (int64_t)blockIdx.x * 64 + lane;
  if (g >= T * 4) return;
  int64_t tk = g >> 2;
  int s = (int)(g & 3);
  uint64_t key = task_key[tk];
  uint32_t ev = (uint32_t)(key >> 13);
  int ch = (int)(key & 8191u) * 4 + s;
  int a = task_off[tk], b = task_off[tk + 1];
  LaneRecs R{c_time, c_v, c_row, c_pix, coff[g], coff[g + 1] - coff[g], 0};
  c_count[g] = 0;
  if (ch >= t.n_channels || (t.status[ch] & 0b101) != 0) return;
  for (int w = 0; w < CRV_DEAD_WORDS; ++w) L.bm[w * 64 + lane] = 0u;
  Draw dr{(uint32_t)p.seed, p.evkey_hi};
  const double t_end = p.photon_window_hi;
  // start the photon streams
  for (int k = a; k < b; ++k) {
    int64_t gi = (int64_t)k * 4 + s;
    int32_t N = nph[gi];
    gen.r[gi] = 0;
    if (N <= 0) continue;
    int32_t i = perm[k];
    uint32_t j = (uint32_t)(i - evfirst[i]);
    double d = dist_to_end(p, st.z[i], s);
    double tb = st.t[i] + d * p.inv_speed;
    double u = unif(dr(KIND_DELAY, ev, (uint32_t)ch, j, 0u).x);
    double S = 0.0 + (-fm_log(u)) / (double)N;
    gen.S[gi] = S;
    gen.tb[gi] = tb;
    gen.head[gi] = tb + p.tau * S;
  }
  bool overflow = false;
  for (;;) {
    // next photon: the earliest stream head (lowest step first on ties)
    int kb = -1;
    double tbest = 0.0;
    for (int k = a; k < b; ++k) {
      int64_t gi = (int64_t)k * 4 + s;
      if (gen.r[gi] >= nph[gi]) continue;
      double h = gen.head[gi];
      if (kb < 0 || h < tbest) {
        kb = k;
        tbest = h;
      }
    }
    if (kb < 0 || tbest > t_end) break;  // later photons of every stream are later still
    int64_t gi = (int64_t)kb * 4 + s;
    int32_t row = perm[kb];
    uint32_t j = (uint32_t)(row - evfirst[row]);
    int32_t r = gen.r[gi];
    int32_t N = nph[gi];
    const double tph = tbest;
    // advance the stream
    if (r + 1 < N) {
      int32_t rn = r + 1;
      U4 uu = dr(KIND_DELAY, ev, (uint32_t)ch, j, (uint32_t)(rn >> 2));
      uint32_t x = (rn & 3) == 0 ? uu.x : (rn & 3) == 1 ? uu.y : (rn & 3) == 2 ? uu.z : uu.w;
      double S = gen.S[gi] + (-fm_log(unif(x))) / (double)(N - rn);
      gen.S[gi] = S;
      gen.head[gi] = gen.tb[gi] + p.tau * S;
    }
    gen.r[gi] = r + 1;
    // pixel and firing of the photon
    U4 u = dr(KIND_PIXEL, ev, (uint32_t)ch, j, pack_photon((uint32_t)r, 0u, 0u));
    int bin = upper_bound(t.cdf, t.ncdf, unif(u.x) * t.cdf_total);
    int px = bin / t.map_ny, py = bin % t.map_ny;
    if (p.pixel_lrint) {
      px += unif(u.y) >= 0.5 ? 1 : 0;
      py += unif(u.z) >= 0.5 ? 1 : 0;
    }
    int pix = px * CRV_NPIX + py;  // px, py <= 40: the photon map is at most 40 x 40 bins
    if (px_dead(t.dead, pix)) continue;
    double v = p.overvoltage * (1.0 - recovery(p, L, lane, R, pix, tph));
    double pf = p.p1 * (1.0 - fm_exp(-v / p.p2));
    double pr = 1.0 - (1.0 - pf);
    if (!(unif(u.w) < pr)) continue;
    if (!add_record(R, tph, v, row, pix)) {
      overflow = true;
      break;
    }
    px_set(L, lane, pix);
    if (!cascade(p, t, dv, dr, ev, (uint32_t)ch, j, (uint32_t)r, tph, row, pix, R, L, lane, flags,
                 [](int, int, double) {})) {
      overflow = true;
      break;
    }
  }
  if (overflow) atomicAdd(&flags[F_OVERFLOW], 1);
  c_count[g] = R.cnt;
}

// photoelectrons of each (hit counter, SiPM) lane: its steps in task order, each step's photons in rank order (the
// ordered delays of k_sipm's streams), at poff[g] onwards; times shifted and folded when `shift` is given
__global__ void k_photons_out(crv_params p, Steps st, int64_t T, const uint64_t *__restrict__ task_key,
                              const int32_t *__restrict__ task_off, const int32_t *__restrict__ perm,
                              const int32_t *__restrict__ evfirst, const int32_t *__restrict__ nph,
                              const int64_t *__restrict__ poff, const double *__restrict__ shift, double period,
                              crv_photons out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= T * 4) return;
  const int64_t tk = g >> 2;
  const int s = (int)(g & 3);
  const uint64_t key = task_key[tk];
  const uint32_t ev = (uint32_t)(key >> 13);
  const int ch = (int)(key & 8191u) * 4 + s;
  const int64_t label = ((((int64_t)p.evkey_hi) << 32) | (int64_t)ev) * CRV_CHAN_MULT + ch;
  const Draw dr{(uint32_t)p.seed, p.evkey_hi};
  int64_t o = poff[g];
  for (int k = task_off[tk]; k < task_off[tk + 1]; ++k) {
    const int32_t N = nph[(int64_t)k * 4 + s];
    if (N <= 0) continue;
    const int32_t i = perm[k];
    const uint32_t j = (uint32_t)(i - evfirst[i]);
    const double tb = st.t[i] + dist_to_end(p, st.z[i], s) * p.inv_speed;
    double S = 0.0;
    U4 uu{0u, 0u, 0u, 0u};
    for (int32_t r = 0; r < N; ++r, ++o) {
      if ((r & 3) == 0) uu = dr(KIND_DELAY, ev, (uint32_t)ch, j, (uint32_t)(r >> 2));
      const uint32_t x = (r & 3) == 0 ? uu.x : (r & 3) == 1 ? uu.y : (r & 3) == 2 ? uu.z : uu.w;
      S = S + (-fm_log(unif(x))) / (double)(N - r);
      double tt = tb + p.tau * S;
      if (shift) {
        tt = fmod(tt + shift[ev], period);
        if (tt < 0) tt += period;
      }
      out.label[o] = label;
      out.time[o] = tt;
      out.photon[o] = i;
      out.j[o] = j;
      out.rank[o] = (uint32_t)r;
    }
  }
}

__global__ void k_charges_out(crv_params p, int64_t T, const uint64_t *__restrict__ task_key,
                              const int64_t *__restrict__ coff, const int64_t *__restrict__ ooff,
                              const double *__restrict__ c_time, const double *__restrict__ c_v,
                              const int32_t *__restrict__ c_row, int64_t *__restrict__ label,
                              double *__restrict__ time, double *__restrict__ charge, double *__restrict__ pe,
                              int64_t *__restrict__ photon) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= T * 4) return;
  uint64_t key = task_key[g >> 2];
  int64_t evkey = ((int64_t)p.evkey_hi << 32) | (int64_t)(key >> 13);
  int64_t lab = evkey * CRV_CHAN_MULT + (int64_t)((key & 8191u) * 4 + (g & 3));
  int64_t src = coff[g], dst = ooff[g], n = ooff[g + 1] - dst;
  for (int64_t q = 0; q < n; ++q) {
    double v = c_v[src + q];
    label[dst + q] = lab;
    time[dst + q] = c_time[src + q];
    charge[dst + q] = p.capacitance * v;
    pe[dst + q] = v / p.overvoltage;
    photon[dst + q] = c_row[src + q];
  }
}

}  // namespace

int group_steps(Ctx *c, const crv_steps *st, StepGroups &g, hipStream_t s) {
  const int64_t n = st->n;
  g.n_steps = n;
  g.ev = st->event;
  g.counter = st->counter;
  g.edep = st->edep;
  g.z = st->z;
  g.t = st->t;
  g.T = 0;
  if (n == 0) return CRV_OK;
  if (n > INT32_MAX / 4) return set_error(c, CRV_E_ARG, "too many steps for one call");
  uint64_t *key = (uint64_t *)ws(c, 0, n * 8);
  uint64_t *key_s = (uint64_t *)ws(c, 1, n * 8);
  int32_t *idx = (int32_t *)ws(c, 2, n * 4);
  g.perm = (int32_t *)ws(c, 3, n * 4);
  g.evfirst = (int32_t *)ws(c, 4, n * 4);
  g.task_key = (uint64_t *)ws(c, 5, n * 8);
  int32_t *cnt = (int32_t *)ws(c, 6, (n + 1) * 4);
  g.task_off = (int32_t *)ws(c, 7, (n + 1) * 4);
  int32_t *nruns = (int32_t *)ws(c, 8, 8);
  if (!key || !key_s || !idx || !g.perm || !g.evfirst || !g.task_key || !cnt || !g.task_off || !nruns)
    return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  k_step_keys<<<nblocks(n, 256), 256, 0, s>>>(st->event, st->counter, n, key, idx, g.evfirst, c->flags);
  CRV_HIP(c, hipGetLastError());
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key_s, idx, g.perm, (int)n, 0, 45, s));
  void *tmp = ws(c, 9, tb);
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key_s, idx, g.perm, (int)n, 0, 45, s));
  size_t tb2 = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, key_s, g.task_key, cnt, nruns, (int)n, s));
  tmp = ws(c, 9, tb2);
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp, tb2, key_s, g.task_key, cnt, nruns, (int)n, s));
  int32_t hr = 0;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(&hr, nruns, 4, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_UNSORTED]) return set_error(c, CRV_E_ARG, "steps must be sorted by event");
  if (hf[F_COUNTER]) return set_error(c, CRV_E_ARG, "counter id >= 8192");
  g.T = hr;
  size_t tb3 = 0;
  CRV_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cnt, g.task_off, hr + 1, s));
  tmp = ws(c, 9, tb3);
  CRV_HIP(c, hipMemsetAsync(cnt + hr, 0, 4, s));
  CRV_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb3, cnt, g.task_off, hr + 1, s));
  return CRV_OK;
}

namespace {

// k_step_keys over the capacity of a fixed-capacity call: steps past the call's count get a key above every real one
// (bit 45), so they sort last and form one run of their own
__global__ void k_step_keys_fixed(const uint32_t *__restrict__ ev, const uint32_t *__restrict__ counter,
                                  const int64_t *__restrict__ n_dev, int64_t n_cap, uint64_t *__restrict__ key,
                                  int32_t *__restrict__ idx, int32_t *__restrict__ evfirst, int *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cap) return;
  const int64_t n = *n_dev;
  idx[i] = (int32_t)i;
  if (i >= n) {
    key[i] = 1ull << 45;
    return;
  }
  const uint32_t e = ev[i];
  int64_t lo = 0, hi = i;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ev[mid] < e) lo = mid + 1; else hi = mid;
  }
  evfirst[i] = (int32_t)lo;
  if (i > 0 && ev[i - 1] > e) atomicAdd(&flags[F_UNSORTED], 1);
  const uint32_t c = counter[i];
  if (c >= 8192u) atomicAdd(&flags[F_COUNTER], 1);
  key[i] = ((uint64_t)e << 13) | (uint64_t)(c & 8191u);
}

// hit counters = runs of real keys (the padding forms one more run when the call is below capacity); the run counts
// past them are cleared so that the task offsets end at the last real step
__global__ void k_fix_tasks(const int32_t *__restrict__ nruns, const int64_t *__restrict__ n_dev, int64_t n_cap,
                            int32_t *__restrict__ T_dev, int32_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = *n_dev;
  const int32_t T = n == 0 ? 0 : *nruns - (n < n_cap ? 1 : 0);
  if (i == 0) *T_dev = T;
  if (i >= T && i <= n_cap) cnt[i] = 0;
}

__global__ void k_zero_flags(int *__restrict__ f, int n) {
  const int i = threadIdx.x;
  if (i < n) f[i] = 0;
}

}  // namespace

int group_steps_fixed(Ctx *c, Plan &pl, Arena &ar, StepGroups &g, hipStream_t s) {
  const int64_t n = pl.n_cap;
  g.n_steps = n;
  g.ev = pl.in_event;
  g.counter = pl.in_counter;
  g.edep = pl.in_edep;
  g.z = pl.in_z;
  g.t = pl.in_t;
  g.T = n;  // capacity; the call's count is in T_dev
  uint64_t *key = (uint64_t *)ar.get(n * 8), *key_s = (uint64_t *)ar.get(n * 8);
  int32_t *idx = (int32_t *)ar.get(n * 4);
  g.perm = (int32_t *)ar.get(n * 4);
  g.evfirst = (int32_t *)ar.get(n * 4);
  g.task_key = (uint64_t *)ar.get(n * 8);
  int32_t *cnt = (int32_t *)ar.get((n + 1) * 4);
  g.task_off = (int32_t *)ar.get((n + 1) * 4);
  int32_t *nruns = (int32_t *)ar.get(8);
  if (!key || !key_s || !idx || !g.perm || !g.evfirst || !g.task_key || !cnt || !g.task_off || !nruns)
    return set_error(c, CRV_E_HIP, "graph: step grouping allocation failed" + ar.note());
  k_zero_flags<<<1, 64, 0, s>>>(c->flags, F_NFLAGS);
  k_step_keys_fixed<<<nblocks(n, 256), 256, 0, s>>>(pl.in_event, pl.in_counter, pl.n_dev, n, key, idx, g.evfirst,
                                                   c->flags);
  CRV_HIP(c, hipGetLastError());
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key_s, idx, g.perm, (int)n, 0, 46, s));
  void *tmp = ar.get(tb);
  if (!tmp) return set_error(c, CRV_E_HIP, "graph: sort buffer allocation failed" + ar.note());
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, key_s, idx, g.perm, (int)n, 0, 46, s));
  size_t tb2 = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, key_s, g.task_key, cnt, nruns, (int)n, s));
  void *tmp2 = ar.get(tb2);
  if (!tmp2) return set_error(c, CRV_E_HIP, "graph: run-length buffer allocation failed" + ar.note());
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp2, tb2, key_s, g.task_key, cnt, nruns, (int)n, s));
  k_fix_tasks<<<nblocks(n + 1, 256), 256, 0, s>>>(nruns, pl.n_dev, n, pl.T_dev, cnt);
  CRV_HIP(c, hipGetLastError());
  size_t tb3 = 0;
  CRV_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cnt, g.task_off, (int)(n + 1), s));
  void *tmp3 = ar.get(tb3);
  if (!tmp3) return set_error(c, CRV_E_HIP, "graph: scan buffer allocation failed" + ar.note());
  CRV_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmp3, tb3, cnt, g.task_off, (int)(n + 1), s));
  return CRV_OK;
}

int photons_stage(Ctx *c, StepGroups &g, const double *shift, double period, crv_photons *out, hipStream_t s) {
  const int64_t T = g.T, L = T * 4;
  out->capped = 0;
  if (shift && !(std::isfinite(period) && period > 0))
    return set_error(c, CRV_E_ARG, "photons: the fold period must be finite and positive");
  if (T == 0) {
    out->n = 0;
    return CRV_OK;
  }
  Steps st{g.ev, g.counter, g.edep, g.z, g.t};
  int32_t *nph = (int32_t *)ws(c, 10, g.n_steps * 4 * 4);
  int64_t *sipm_nph = (int64_t *)ws(c, 11, (L + 1) * 8);
  int64_t *poff = (int64_t *)ws(c, 13, (L + 1) * 8);
  if (!nph || !sipm_nph || !poff) return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  k_photon_count<<<nblocks(L, 256), 256, 0, s>>>(c->p, c->t, c->d, st, T, g.task_key, g.task_off, g.perm, g.evfirst,
                                                 nph, sipm_nph, c->flags);
  CRV_HIP(c, hipGetLastError());
  int rc = scan_i64(c, sipm_nph, poff, L, 20, s);
  if (rc) return rc;
  int64_t nphot = 0;
  if ((rc = read_i64(c, poff + L, &nphot, s))) return rc;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "channel beyond the uploaded conditions");
  if (hf[F_OVERFLOW]) return set_error(c, CRV_E_OVERFLOW, "photons: more than 65535 photoelectrons from one step");
  out->capped = hf[F_CAPPED];
  out->n = nphot;
  if (nphot > out->capacity) return set_error(c, CRV_E_CAPACITY, "photons: output capacity too small");
  k_photons_out<<<nblocks(L, 256), 256, 0, s>>>(c->p, st, T, g.task_key, g.task_off, g.perm, g.evfirst, nph, poff,
                                                shift, period, *out);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

int charges_stage(Ctx *c, StepGroups &g, crv_charges *out, hipStream_t s) {
  const int64_t T = g.T, L = T * 4;
  out->n_photons = 0;
  out->capped = 0;
  if (T == 0) {
    out->n = 0;
    return CRV_OK;
  }
  Steps st{g.ev, g.counter, g.edep, g.z, g.t};
  int32_t *nph = (int32_t *)ws(c, 10, g.n_steps * 4 * 4);
  int64_t *sipm_nph = (int64_t *)ws(c, 11, (L + 1) * 8);
  int64_t *cap = (int64_t *)ws(c, 12, (L + 1) * 8);
  int64_t *coff = (int64_t *)ws(c, 13, (L + 1) * 8);
  int32_t *ccount = (int32_t *)ws(c, 14, (L + 1) * 4);
  int64_t *ooff = (int64_t *)ws(c, 15, (L + 1) * 8);
  Gen gen;
  gen.S = (double *)ws(c, 16, g.n_steps * 4 * 8);
  gen.head = (double *)ws(c, 17, g.n_steps * 4 * 8);
  gen.tb = (double *)ws(c, 18, g.n_steps * 4 * 8);
  gen.r = (int32_t *)ws(c, 19, g.n_steps * 4 * 4);
  if (!nph || !sipm_nph || !cap || !coff || !ccount || !ooff || !gen.S || !gen.head || !gen.tb || !gen.r)
    return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  k_photon_count<<<nblocks(L, 256), 256, 0, s>>>(c->p, c->t, c->d, st, T, g.task_key, g.task_off, g.perm, g.evfirst,
                                                 nph, sipm_nph, c->flags);
  CRV_HIP(c, hipGetLastError());
  int rc = scan_i64(c, sipm_nph, coff, L, 20, s);  // coff used as scratch for the photon total
  if (rc) return rc;
  int64_t nphot = 0;
  if ((rc = read_i64(c, coff + L, &nphot, s))) return rc;
  out->n_photons = nphot;
  for (int factor = 2; factor <= 64; factor *= 4) {
    k_charge_caps<<<nblocks(L, 256), 256, 0, s>>>(sipm_nph, L, cap, factor);
    if ((rc = scan_i64(c, cap, coff, L, 21, s))) return rc;
    int64_t total = 0;
    if ((rc = read_i64(c, coff + L, &total, s))) return rc;
    double *c_time = (double *)ws(c, 22, total * 8 + 8);
    double *c_v = (double *)ws(c, 23, total * 8 + 8);
    int32_t *c_row = (int32_t *)ws(c, 24, total * 4 + 8);
    int16_t *c_pix = (int16_t *)ws(c, 25, total * 2 + 8);
    if (!c_time || !c_v || !c_row || !c_pix) return set_error(c, CRV_E_HIP, "charge workspace allocation failed");
    CRV_HIP(c, hipMemsetAsync(c->flags + F_OVERFLOW, 0, sizeof(int) * 2, s));
    kmark(c, 0, 0, s);
    k_sipm<<<nblocks(L, 64), 64, 0, s>>>(c->p, c->t, c->d, st, T, g.task_key, g.task_off, g.perm, g.evfirst, nph,
                                         coff, gen, c_time, c_v, c_row, c_pix, ccount, c->flags);
    CRV_HIP(c, hipGetLastError());
    kmark(c, 0, 1, s);
    int hf[F_NFLAGS];
    CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    CRV_HIP(c, hipStreamSynchronize(s));
    out->capped = hf[F_CAPPED];
    if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "channel beyond the uploaded conditions");
    if (hf[F_CASCADE]) return set_error(c, CRV_E_OVERFLOW, "SiPM cross-talk cascade deeper than 31 generations");
    if (hf[F_OVERFLOW]) continue;  // more room per SiPM, then again
    if ((rc = scan_i32_to_i64(c, ccount, ooff, L, 26, s))) return rc;
    int64_t nq = 0;
    if ((rc = read_i64(c, ooff + L, &nq, s))) return rc;
    out->n = nq;
    if (out->capacity < 0) {  // internal outputs of the chained run
      out->label = (int64_t *)ws(c, 300, nq * 8 + 8);
      out->time = (double *)ws(c, 301, nq * 8 + 8);
      out->charge = (double *)ws(c, 302, nq * 8 + 8);
      out->pe = (double *)ws(c, 303, nq * 8 + 8);
      out->photon = (int64_t *)ws(c, 304, nq * 8 + 8);
      if (!out->label || !out->time || !out->charge || !out->pe || !out->photon)
        return set_error(c, CRV_E_HIP, "charge table allocation failed");
      out->capacity = nq;
    }
    if (nq > out->capacity) return set_error(c, CRV_E_CAPACITY, "charges: output capacity too small");
    k_charges_out<<<nblocks(L, 256), 256, 0, s>>>(c->p, T, g.task_key, coff, ooff, c_time, c_v, c_row, out->label,
                                                  out->time, out->charge, out->pe, out->photon);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipStreamSynchronize(s));
    return CRV_OK;
  }
  return set_error(c, CRV_E_OVERFLOW, "SiPM charges exceed 17 per photon or a cascade generation exceeds 64 pixels");
}

}  // namespace crv
