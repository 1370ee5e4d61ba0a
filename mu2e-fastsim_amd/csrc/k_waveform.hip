// Digitisation stage: SiPM charges -> 128-sample ADC records and zero-suppressed digis (FP64).
//
// Follows the reference's reco/waveform.py:65-172 (charge clusters split at gaps > 100 ns + 4 periods, first FEB
// tick after first charge - 12.5 ns, template sums in charge order, Gaussian noise, zero suppression on samples
// i-2..i+2 > 27.5 mV, 12-sample digis, ADC = int(V * 2300 + 100 + 0.5) clipped) and chain.py:219-231 (one TDC0 per
// (event, ROC, FEB), here a Philox draw keyed by them). One lane per SiPM label walks its time-ordered charges once
// per cluster: the template contributions of a charge reach samples [j0, j0 + 10) and j0 never decreases, so a
// 16-sample window in LDS holds every open sample; a sample is final once the next charge starts past it, then it
// gets its noise, its ADC word in the record and its zero-suppression bit (decided two samples later).
#include <hipcub/hipcub.hpp>

#include "crv_internal.h"

namespace crv {

namespace {

constexpr int RING = 16;
constexpr double PERIOD = 12.5;
constexpr double GAP_PERIODS = 4.0;

__device__ __forceinline__ int adc_of(const crv_params &p, double v) {
  double x = (v * p.adc_conversion + (double)p.pedestal) + 0.5;
  x = fmin(fmax(x, -1.0), (double)p.max_adc + 1.0);
  double a = trunc(x);
  a = fmin(fmax(a, (double)p.min_adc), (double)p.max_adc);
  return (int)a;
}

__global__ void k_wf_count(crv_params p, int64_t nseg, const int64_t *__restrict__ soff,
                           const double *__restrict__ time, int32_t *__restrict__ nclus) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const double gap = p.single_pe_max_time + GAP_PERIODS * PERIOD;
  int64_t a = soff[g], b = soff[g + 1];
  int32_t n = 1;
  for (int64_t q = a + 1; q < b; ++q)
    if (time[q] - time[q - 1] > gap) ++n;
  nclus[g] = n;
}

struct DigiSlot {
  int32_t *rec;     // record slot
  int32_t *sample;  // first sample
  double *start;
};

__global__ __launch_bounds__(64) void k_wf_synth(crv_params p, Tables t, int32_t nreach, int64_t nseg,
                                                 const int64_t *__restrict__ seg_label, const int64_t *__restrict__ soff,
                                                 const double *__restrict__ time, const double *__restrict__ charge,
                                                 const int64_t *__restrict__ roff, const int64_t *__restrict__ doff,
                                                 uint16_t *__restrict__ rec_adc, uint8_t *__restrict__ rec_has,
                                                 int32_t *__restrict__ rec_tick, double *__restrict__ rec_tdc0,
                                                 DigiSlot ds, int32_t *__restrict__ ndig, int *__restrict__ flags) {
  __shared__ double ring[RING * 64];
  const int lane = threadIdx.x;
  int64_t g = (int64_t)blockIdx.x * 64 + lane;
  if (g >= nseg) return;
  const int W = p.window_samples, Sz = p.samples_zs;
  const int lmax = W - Sz;
  const double gap = p.single_pe_max_time + GAP_PERIODS * PERIOD;
  int64_t label = seg_label[g];
  int64_t evkey = label / CRV_CHAN_MULT;
  int ch = (int)(label % CRV_CHAN_MULT);
  ndig[g] = 0;
  if (ch < 0 || ch >= t.n_channels) {
    atomicAdd(&flags[F_BADCHAN], 1);
    for (int64_t rs = roff[g]; rs < roff[g + 1]; ++rs) rec_has[rs] = 0;
    return;
  }
  Draw dr{(uint32_t)p.seed, p.evkey_hi};
  uint32_t fk = p.feb_key_roc ? (uint32_t)(t.roc[ch] * 1000 + t.feb[ch]) : (uint32_t)t.feb[ch];
  const double td = unif(dr(KIND_TDC0, (uint32_t)evkey, fk, 0u, 0u).x) * p.feb_phase_width;
  int64_t a = soff[g], b = soff[g + 1];
  int64_t rslot = roff[g];
  int64_t dbase = doff[g], dcap = doff[g + 1] - dbase;
  int32_t nd = 0;
  int32_t cl = 0;
  int overlong = 0;
  for (int64_t q0 = a; q0 < b; ++cl) {
    int64_t q1 = q0 + 1;
    while (q1 < b && !(time[q1] - time[q1 - 1] > gap)) ++q1;
    const int64_t rs = rslot + cl;
    rec_has[rs] = 0;
    const double nq = (double)(q1 - q0);
    const bool keep = nq * t.vmax >= p.min_voltage;
    const double t_first = time[q0] - PERIOD;
    const double kf = ceil((t_first - td) / PERIOD);
    const double t_start = kf * PERIOD + td;
    rec_tick[rs] = (int32_t)kf;
    rec_tdc0[rs] = td;
    if (!keep) {
      q0 = q1;
      continue;
    }
    const int64_t est = (int64_t)floor((time[q1 - 1] - time[q0] + p.single_pe_max_time) / PERIOD);
    // pass A: the last sample any charge reaches (with and without the window limit)
    int64_t used = 0, used_full = 0;
    for (int64_t q = q0; q < q1; ++q) {
      const double tq = time[q];
      int64_t j0 = (int64_t)fmax(ceil((tq - t_start) / PERIOD), 0.0);
      for (int r = 0; r < nreach; ++r) {
        int64_t J = j0 + r;
        double st = (double)J * PERIOD + t_start - tq;
        double ti = rint(st / p.single_pe_precision);
        if (!(st >= 0.0) || !(ti < (double)t.ntmpl)) continue;
        if (J + 1 > used_full) used_full = J + 1;
        if (J < lmax && J + 1 > used) used = J + 1;
      }
    }
    int64_t length = est > used ? est : used;
    if ((est > used_full ? est : used_full) > lmax) ++overlong;
    if (length > lmax) length = lmax;
    // pass B: sliding sums, noise, record ADC, zero suppression
    for (int k = 0; k < RING; ++k) ring[k * 64 + lane] = 0.0;
    uint32_t above = 0u;  // bit k: sample E - 1 - k above threshold
    int64_t E = 0;        // samples < E are final
    int64_t next_i = 0;
    uint16_t *rec = rec_adc + rs * (int64_t)W;
    auto finalize = [&](int64_t J) {
      double w = ring[(J % RING) * 64 + lane];
      ring[(J % RING) * 64 + lane] = 0.0;
      bool valid = J < length;
      if (valid && p.noise_on) {
        U4 u = dr(KIND_NOISE, (uint32_t)evkey, (uint32_t)ch, (uint32_t)(int32_t)kf, (uint32_t)(J >> 1));
        double z = (J & 1) ? box_muller(unif(u.z), unif(u.w)) : box_muller(unif(u.x), unif(u.y));
        w = w + p.noise * z;
      }
      if (J < W) rec[J] = (uint16_t)adc_of(p, valid ? w : 0.0);
      bool ab = valid && (w > p.min_voltage);
      above = (above << 1) | (ab ? 1u : 0u);
      // zero suppression for i = J - 2 (needs above[i - 2 .. i + 2] = bits 4..0)
      int64_t i = J - 2;
      if (i >= 0 && i < length && (above & 0x1Fu) != 0u && next_i <= i) {
        double tstart = t_start + (double)i * PERIOD;
        if (tstart >= p.window_lo && tstart <= p.window_hi) {
          if (nd < dcap) {
            ds.rec[dbase + nd] = (int32_t)rs;
            ds.sample[dbase + nd] = (int32_t)i;
            ds.start[dbase + nd] = tstart;
          }
          ++nd;
          next_i = i + Sz;
        }
      }
    };
    for (int64_t q = q0; q < q1; ++q) {
      const double tq = time[q];
      int64_t j0 = (int64_t)fmax(ceil((tq - t_start) / PERIOD), 0.0);
      while (E < j0 && E < length + 2) finalize(E++);
      const double w = charge[q] / p.single_pe_reference_charge;
      for (int r = 0; r < nreach; ++r) {
        int64_t J = j0 + r;
        double st = (double)J * PERIOD + t_start - tq;
        double ti = rint(st / p.single_pe_precision);
        if (!(st >= 0.0) || !(ti < (double)t.ntmpl) || J >= lmax) continue;
        ring[(J % RING) * 64 + lane] += t.tmpl[(int)ti] * w;
      }
    }
    while (E < length + 2) finalize(E++);
    for (int64_t J = E; J < W; ++J) rec[J] = (uint16_t)adc_of(p, 0.0);
    if (nd > 0 && (nd > dcap || ds.rec[dbase + nd - 1] == (int32_t)rs)) rec_has[rs] = 1;
    q0 = q1;
  }
  if (overlong) atomicAdd(&flags[F_OVERLONG], overlong);
  if (nd > dcap) {
    atomicAdd(&flags[F_OVERFLOW], 1);
    nd = (int32_t)dcap;
  }
  // order the digis of the label by (start, sample, cluster) (the reference's lexsort over its emission order)
  for (int32_t x = 1; x < nd; ++x) {
    int32_t rr = ds.rec[dbase + x], ss = ds.sample[dbase + x];
    double st = ds.start[dbase + x];
    int32_t y = x - 1;
    while (y >= 0) {
      double sy = ds.start[dbase + y];
      int32_t iy = ds.sample[dbase + y], ry = ds.rec[dbase + y];
      bool greater = sy > st || (sy == st && (iy > ss || (iy == ss && ry > rr)));
      if (!greater) break;
      ds.rec[dbase + y + 1] = ry;
      ds.sample[dbase + y + 1] = iy;
      ds.start[dbase + y + 1] = sy;
      --y;
    }
    ds.rec[dbase + y + 1] = rr;
    ds.sample[dbase + y + 1] = ss;
    ds.start[dbase + y + 1] = st;
  }
  ndig[g] = nd;
}

__global__ void k_scale(const int64_t *__restrict__ in, int64_t *__restrict__ out, int64_t n, int f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * f;
}

__global__ void k_rec_flags(int64_t nrec, const uint8_t *__restrict__ has, int64_t *__restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrec) f[i] = has[i];
}

__global__ void k_wf_out(crv_params p, int64_t nseg, const int64_t *__restrict__ seg_label,
                         const int64_t *__restrict__ roff, const int64_t *__restrict__ doff,
                         const int64_t *__restrict__ dout, const int64_t *__restrict__ rout,
                         const uint16_t *__restrict__ rec_adc, const uint8_t *__restrict__ rec_has,
                         const int32_t *__restrict__ rec_tick, const double *__restrict__ rec_tdc0, DigiSlot ds,
                         crv_digis o) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const int W = p.window_samples, Sz = p.samples_zs;
  int64_t label = seg_label[g];
  for (int64_t rs = roff[g]; rs < roff[g + 1]; ++rs) {
    if (!rec_has[rs]) continue;
    int64_t ro = rout[rs];
    if (o.record_adc)
      for (int k = 0; k < W; ++k) o.record_adc[ro * W + k] = rec_adc[rs * W + k];
    if (o.record_label) o.record_label[ro] = label;
    if (o.record_tick) o.record_tick[ro] = rec_tick[rs];
  }
  int64_t d0 = dout[g], n = dout[g + 1] - d0, dbase = doff[g];
  for (int64_t k = 0; k < n; ++k) {
    int32_t rs = ds.rec[dbase + k];
    int32_t i = ds.sample[dbase + k];
    double start = ds.start[dbase + k];
    double td = rec_tdc0[rs];
    int64_t o_ = d0 + k;
    o.label[o_] = label;
    o.start[o_] = start;
    o.tdc0[o_] = td;
    o.starttdc[o_] = (int64_t)rint((start - td) / PERIOD);
    if (o.record) o.record[o_] = (int32_t)rout[rs];
    if (o.sample) o.sample[o_] = i;
    for (int s = 0; s < Sz; ++s) o.adc[o_ * Sz + s] = (int16_t)rec_adc[(int64_t)rs * W + i + s];
  }
}

}  // namespace

int digis_stage(Ctx *c, const int64_t *label, const double *time, const double *charge, int64_t n, crv_digis *out,
                hipStream_t s) {
  const crv_params &p = c->p;
  out->overlong = 0;
  if (p.window_samples < p.samples_zs + 4 || p.window_samples > 4096)
    return set_error(c, CRV_E_ARG, "window_samples must be in [samples_zs + 4, 4096]");
  if (n == 0) {
    out->n = 0;
    out->n_records = 0;
    return CRV_OK;
  }
  int64_t *seg_label = (int64_t *)ws(c, 30, n * 8);
  int32_t *seg_cnt = (int32_t *)ws(c, 31, (n + 1) * 4);
  int32_t *nruns = (int32_t *)ws(c, 32, 8);
  int64_t *soff = (int64_t *)ws(c, 33, (n + 1) * 8);
  if (!seg_label || !seg_cnt || !nruns || !soff) return ws_fail(c);
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, label, seg_label, seg_cnt, nruns, (int)n, s));
  void *tmp = ws(c, 34, tb);
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp, tb, label, seg_label, seg_cnt, nruns, (int)n, s));
  int32_t hs = 0;
  CRV_HIP(c, hipMemcpyAsync(&hs, nruns, 4, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  const int64_t nseg = hs;
  int rc = scan_i32_to_i64(c, seg_cnt, soff, nseg, 35, s);
  if (rc) return rc;
  int32_t *nclus = (int32_t *)ws(c, 36, (nseg + 1) * 4);
  int64_t *roff = (int64_t *)ws(c, 37, (nseg + 1) * 8);
  int64_t *doff = (int64_t *)ws(c, 39, (nseg + 1) * 8);
  k_wf_count<<<nblocks(nseg, 256), 256, 0, s>>>(p, nseg, soff, time, nclus);
  CRV_HIP(c, hipGetLastError());
  if ((rc = scan_i32_to_i64(c, nclus, roff, nseg, 40, s))) return rc;
  int64_t nrec = 0;
  if ((rc = read_i64(c, roff + nseg, &nrec, s))) return rc;
  const int dmax = (p.window_samples - p.samples_zs + p.samples_zs - 1) / p.samples_zs;
  // digi capacity per label: dmax per cluster
  k_scale<<<nblocks(nseg + 1, 256), 256, 0, s>>>(roff, doff, nseg + 1, dmax);
  CRV_HIP(c, hipGetLastError());
  const int W = p.window_samples;
  uint16_t *rec_adc = (uint16_t *)ws(c, 43, (nrec + 1) * W * 2);
  uint8_t *rec_has = (uint8_t *)ws(c, 44, nrec + 8);
  int32_t *rec_tick = (int32_t *)ws(c, 45, (nrec + 1) * 4);
  double *rec_tdc0 = (double *)ws(c, 46, (nrec + 1) * 8);
  DigiSlot ds;
  ds.rec = (int32_t *)ws(c, 47, (nrec * dmax + 1) * 4);
  ds.sample = (int32_t *)ws(c, 48, (nrec * dmax + 1) * 4);
  ds.start = (double *)ws(c, 49, (nrec * dmax + 1) * 8);
  int32_t *ndig = (int32_t *)ws(c, 50, (nseg + 1) * 4);
  int64_t *dout = (int64_t *)ws(c, 51, (nseg + 1) * 8);
  int64_t *rflag = (int64_t *)ws(c, 52, (nrec + 1) * 8);
  int64_t *rout = (int64_t *)ws(c, 53, (nrec + 1) * 8);
  if (!rec_adc || !rec_has || !rec_tick || !rec_tdc0 || !ds.rec || !ds.sample || !ds.start || !ndig || !dout ||
      !rflag || !rout)
    return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  kmark(c, 1, 0, s);
  k_wf_synth<<<nblocks(nseg, 64), 64, 0, s>>>(p, c->t, c->d.nreach, nseg, seg_label, soff, time, charge, roff, doff,
                                              rec_adc, rec_has, rec_tick, rec_tdc0, ds, ndig, c->flags);
  CRV_HIP(c, hipGetLastError());
  kmark(c, 1, 1, s);
  k_rec_flags<<<nblocks(nrec, 256), 256, 0, s>>>(nrec, rec_has, rflag);
  if ((rc = scan_i64(c, rflag, rout, nrec, 54, s))) return rc;
  if ((rc = scan_i32_to_i64(c, ndig, dout, nseg, 55, s))) return rc;
  int64_t nd = 0, nr = 0;
  if ((rc = read_i64(c, dout + nseg, &nd, s))) return rc;
  if ((rc = read_i64(c, rout + nrec, &nr, s))) return rc;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_OVERFLOW]) return set_error(c, CRV_E_OVERFLOW, "digis: more digis than a cluster window holds");
  if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "digis: channel beyond the uploaded conditions");
  out->overlong = hf[F_OVERLONG];
  out->n = nd;
  out->n_records = nr;
  if (out->capacity < 0) {  // internal outputs of the chained run (records in the workspace too)
    out->label = (int64_t *)ws(c, 305, nd * 8 + 8);
    out->start = (double *)ws(c, 306, nd * 8 + 8);
    out->tdc0 = (double *)ws(c, 307, nd * 8 + 8);
    out->starttdc = (int64_t *)ws(c, 308, nd * 8 + 8);
    out->adc = (int16_t *)ws(c, 309, nd * p.samples_zs * 2 + 8);
    out->record = nullptr;
    out->sample = nullptr;
    out->record_adc = (uint16_t *)ws(c, 360, nr * p.window_samples * 2 + 8);
    out->record_label = (int64_t *)ws(c, 361, nr * 8 + 8);
    out->record_tick = (int32_t *)ws(c, 362, nr * 4 + 8);
    if (!out->label || !out->start || !out->tdc0 || !out->starttdc || !out->adc || !out->record_adc ||
        !out->record_label || !out->record_tick)
      return set_error(c, CRV_E_HIP, "digi table allocation failed");
    out->capacity = nd;
    out->record_capacity = nr;
  }
  if (nd > out->capacity || (out->record_adc && nr > out->record_capacity))
    return set_error(c, CRV_E_CAPACITY, "digis: output capacity too small");
  k_wf_out<<<nblocks(nseg, 256), 256, 0, s>>>(p, nseg, seg_label, roff, doff, dout, rout, rec_adc, rec_has, rec_tick,
                                              rec_tdc0, ds, *out);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

}  // namespace crv
