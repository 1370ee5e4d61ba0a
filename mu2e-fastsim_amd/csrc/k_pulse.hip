// Pulse reconstruction: reco waveforms -> pulses (FP64), one lane per waveform.
//
// Follows the reference's reco/pulsefit.py:236-316 (reconstruct / reconstruct_full) and reco/pulsefit_cpu.py:5-57:
// pedestal rule (first sample when it undershoots the DB pedestal by > 6 ADC), FindNextPeak (rising edge, plateau,
// falling edge above 40 ADC), fit window of up to 4 / 3 monotonic samples trimmed by one, Gumbel fit
// A exp(-z - e^{-z}) in Minuit's sine-mapped coordinates with 40 Levenberg-Marquardt iterations and 8 Gauss-Newton
// steps halved up to three times, EDM and Hessian checks, 1 % limit proximity, peak-sample fallback for failed fits,
// subtraction of the rounded fit and the next round from the fit end; then filter_channels and the DB time offset
// (chain.py:317-321). The normal equations are accumulated point by point in window order and solved by Gaussian
// elimination with partial pivoting; eigenvalues by six cyclic Jacobi sweeps.
#include <hipcub/hipcub.hpp>

#include <map>
#include <mutex>

#include "crv_internal.h"

namespace crv {

namespace {

constexpr double PERIOD = 12.5;

struct Sums {
  double chi2;
  double jtj[6];  // 00 01 02 11 12 22
  double jtr[3];
};

struct FitIn {
  const double *y;  // waveform samples (pedestal-subtracted)
  int fs, fe;
  int64_t start_tdc;
  double lo[3], hi[3];
};

__device__ __forceinline__ void params_of(const FitIn &f, const double u[3], double p[3], double dpdu[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double span = f.hi[i] - f.lo[i];
    p[i] = f.lo[i] + span * (fm_sin(u[i]) + 1.0) * 0.5;
    dpdu[i] = span * 0.5 * fm_cos(u[i]);
  }
}

__device__ void eval(const FitIn &f, const double u[3], Sums &s) {
  double p[3], dpdu[3];
  params_of(f, u, p, dpdu);
  const double A = p[0], mu = p[1], b = p[2];
  s.chi2 = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.jtj[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) s.jtr[i] = 0.0;
  for (int q = f.fs; q <= f.fe; ++q) {
    const double t = (double)(f.start_tdc + q) * PERIOD;
    const double z = (t - mu) / b;
    const double ez = fm_exp(-z);
    const double g = fm_exp(-z - ez);
    const double fv = A * g;
    const double r = f.y[q] - fv;
    const double dfdz = fv * (ez - 1.0);
    const double dfdmu = dfdz * (-1.0 / b);
    const double dfdb = dfdz * (-z / b);
    const double J0 = g * dpdu[0], J1 = dfdmu * dpdu[1], J2 = dfdb * dpdu[2];
    s.chi2 = s.chi2 + r * r;
    s.jtr[0] = s.jtr[0] + J0 * r;
    s.jtr[1] = s.jtr[1] + J1 * r;
    s.jtr[2] = s.jtr[2] + J2 * r;
    s.jtj[0] = s.jtj[0] + J0 * J0;
    s.jtj[1] = s.jtj[1] + J0 * J1;
    s.jtj[2] = s.jtj[2] + J0 * J2;
    s.jtj[3] = s.jtj[3] + J1 * J1;
    s.jtj[4] = s.jtj[4] + J1 * J2;
    s.jtj[5] = s.jtj[5] + J2 * J2;
  }
}

// x with (S + diag(add)) x = b, S symmetric given as 6 entries; returns false when a pivot is zero
__device__ bool solve3(const double jtj[6], const double add[3], const double bvec[3], double x[3]) {
  double a[3][3] = {{jtj[0] + add[0], jtj[1], jtj[2]}, {jtj[1], jtj[3] + add[1], jtj[4]}, {jtj[2], jtj[4], jtj[5] + add[2]}};
  double v[3] = {bvec[0], bvec[1], bvec[2]};
  bool sing = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int r = c + 1; r < 3; ++r)
      if (fabs(a[r][c]) > best) {
        piv = r;
        best = fabs(a[r][c]);
      }
    if (piv != c) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double tmp = a[c][k];
        a[c][k] = a[piv][k];
        a[piv][k] = tmp;
      }
      double tv = v[c];
      v[c] = v[piv];
      v[piv] = tv;
    }
    if (a[c][c] == 0.0) sing = true;
    double dd = a[c][c] == 0.0 ? 1.0 : a[c][c];
#pragma unroll
    for (int r = c + 1; r < 3; ++r) {
      double l = a[r][c] / dd;
#pragma unroll
      for (int cc = c + 1; cc < 3; ++cc) a[r][cc] = a[r][cc] - l * a[c][cc];
      v[r] = v[r] - l * v[c];
    }
  }
  double d2 = sing ? 1.0 : a[2][2], d1 = sing ? 1.0 : a[1][1], d0 = sing ? 1.0 : a[0][0];
  x[2] = v[2] / d2;
  x[1] = (v[1] - a[1][2] * x[2]) / d1;
  x[0] = ((v[0] - a[0][1] * x[1]) - a[0][2] * x[2]) / d0;
  return !sing;
}

__device__ void eig3_minmax(const double jtj[6], double &emin, double &emax) {
  double a[3][3] = {{jtj[0], jtj[1], jtj[2]}, {jtj[1], jtj[3], jtj[4]}, {jtj[2], jtj[4], jtj[5]}};
  const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = P[k], q = Q[k], r = 3 - p - q;
      double apq = a[p][q];
      bool nz = apq != 0.0;
      double den = nz ? 2.0 * apq : 1.0;
      double th = (a[q][q] - a[p][p]) / den;
      double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
      t = nz ? t : 0.0;
      double c = 1.0 / sqrt(t * t + 1.0);
      double s = t * c;
      double app = a[p][p] - t * apq;
      double aqq = a[q][q] + t * apq;
      double arp = a[r][p], arq = a[r][q];
      double nrp = c * arp - s * arq;
      double nrq = s * arp + c * arq;
      a[p][p] = app;
      a[q][q] = aqq;
      a[p][q] = a[q][p] = 0.0;
      a[r][p] = a[p][r] = nrp;
      a[r][q] = a[q][r] = nrq;
    }
  }
  double d0 = a[0][0], d1 = a[1][1], d2 = a[2][2];
  emin = d0;
  if (d1 < emin) emin = d1;
  if (d2 < emin) emin = d2;
  emax = d0;
  if (d1 > emax) emax = d1;
  if (d2 > emax) emax = d2;
}

struct FitOut {
  double par[3], chi2;
  bool conv, hess_ok;
};

__device__ void fit_gumbel(const crv_params &pp, const FitIn &f, const double p0[3], FitOut &o) {
  double u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double x = 2.0 * (p0[i] - f.lo[i]) / (f.hi[i] - f.lo[i]) - 1.0;
    x = fmin(fmax(x, -1.0), 1.0);
    u[i] = fm_asin(x);
  }
  double lam = 1e-3;
  Sums s, sn;
  eval(f, u, s);
  for (int it = 0; it < pp.lm_iterations; ++it) {
    const double add[3] = {lam * (s.jtj[0] + 1e-12), lam * (s.jtj[3] + 1e-12), lam * (s.jtj[5] + 1e-12)};
    double du[3];
    if (!solve3(s.jtj, add, s.jtr, du)) du[0] = du[1] = du[2] = 0.0;
    double un[3] = {u[0] + du[0], u[1] + du[1], u[2] + du[2]};
    eval(f, un, sn);
    bool better = sn.chi2 < s.chi2;
    if (better) {
      u[0] = un[0];
      u[1] = un[1];
      u[2] = un[2];
      s = sn;
    } else if (lam >= 1e9) {
      break;  // u, s and lam stay as they are: every remaining iteration would repeat this one
    }
    lam = fmin(fmax(better ? lam * 0.3 : lam * 10.0, 1e-9), 1e9);
  }
  bool stopped = false;
  const double tiny[3] = {1e-12, 1e-12, 1e-12};
  for (int it = 0; it < 8; ++it) {
    double du[3];
    if (!solve3(s.jtj, tiny, s.jtr, du)) stopped = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) du[i] = (isfinite(du[i]) && !stopped) ? du[i] : 0.0;
    const double scales[4] = {1.0, 0.5, 0.25, 0.125};
    bool improved = false;
    for (int k = 0; k < 4; ++k) {
      if (stopped || (du[0] == 0.0 && du[1] == 0.0 && du[2] == 0.0)) break;  // u + 0 cannot lower chi2
      double un[3] = {u[0] + scales[k] * du[0], u[1] + scales[k] * du[1], u[2] + scales[k] * du[2]};
      eval(f, un, sn);
      if (sn.chi2 < s.chi2) {
        u[0] = un[0];
        u[1] = un[1];
        u[2] = un[2];
        s = sn;
        du[0] = du[1] = du[2] = 0.0;
        improved = true;
      }
    }
    if (!improved) break;  // u and s unchanged: the remaining iterations would repeat this one
  }
  double step[3];
  bool ok = solve3(s.jtj, tiny, s.jtr, step);
  double edm = 0.5 * ((s.jtr[0] * step[0] + s.jtr[1] * step[1]) + s.jtr[2] * step[2]);
  double emin, emax;
  eig3_minmax(s.jtj, emin, emax);
  bool hess_ok = emin > 1e-12 * fmax(emax, 1e-300);
  if (!ok) {
    edm = INFINITY;
    hess_ok = false;
  }
  o.conv = isfinite(edm) && edm <= 1e-3 * fmax(s.chi2, 1.0);
  o.hess_ok = hess_ok;
  o.chi2 = s.chi2;
  double dp[3];
  params_of(f, u, o.par, dp);
}

__device__ __forceinline__ double round_half_away(double x) {
  double sg = x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == x ? 0.0 : x));
  return sg * floor(fabs(x) + 0.5);
}

struct PulseSlots {
  int64_t *seq;
  double *pe, *peph, *time, *height, *beta, *chi2, *letime, *ped, *fa, *fmu, *fb;
  uint8_t *failed, *zerondf, *pedfromdb;
};

__global__ void k_pulses(crv_params p, Tables t, int64_t nwf, const int32_t *__restrict__ chan,
                         const int64_t *__restrict__ start, const int32_t *__restrict__ nsamp,
                         const int64_t *__restrict__ off, const int16_t *__restrict__ adc,
                         const double *__restrict__ ped_db, const double *__restrict__ area,
                         const double *__restrict__ height, const double *__restrict__ toff, double *__restrict__ yscr,
                         const int64_t *__restrict__ poff, PulseSlots o, int32_t *__restrict__ npul) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwf) return;
  const int n = nsamp[w];
  const int64_t o0 = off[w];
  const int64_t st = start[w];
  double *y = yscr + o0;
  int32_t np = 0;
  npul[w] = 0;
  if (n <= 0) return;
  const double a0 = (double)adc[o0];
  const bool und = ped_db[w] - a0 > p.pedestal_undershoot_threshold;
  const double ped = und ? a0 : ped_db[w];
  for (int i = 0; i < n; ++i) y[i] = (double)adc[o0 + i] - ped;
  const bool drop = p.ignore_channels && chan && chan[w] >= 0 && chan[w] < t.n_channels &&
                    (t.status[chan[w]] & 0b010) != 0;
  const int64_t pb = poff[w], pcap = poff[w + 1] - pb;
  int startb = 0;
  bool failed_before = false;
  for (int seq = 0;; ++seq) {
    // FindNextPeak
    bool found = false, rising = false;
    int ps = 0, pe = 0;
    for (int b = startb + 1; b < n; ++b) {
      double prev = y[b - 1], cur = y[b];
      if (prev < cur) {
        rising = true;
        ps = b;
        pe = b;
      } else if (prev == cur && rising) {
        pe = b;
      } else if (prev > cur && rising) {
        if (y[pe] > p.min_adc_difference) {
          found = true;
          break;
        }
        rising = false;
      }
    }
    if (!found) break;
    int fs = ps - 1, fe = pe + 1;
    bool go = true;
    for (int k = 1; k <= p.samples_before && go; ++k) {
      int i = ps - k;
      bool okk = i >= 0 && i >= ps - p.samples_before;
      bool cond = okk && y[i] <= y[i + 1];
      if (cond) fs = i;
      go = cond && i != 0;
    }
    go = true;
    for (int k = 1; k <= p.samples_after && go; ++k) {
      int i = pe + k;
      bool cond = i < n && y[i - 1] >= y[i];
      if (cond) fe = i;
      go = cond;
    }
    if (ps - fs > 1) fs += 1;
    if (fe - pe > 1) fe -= 1;
    const double yps = y[ps];
    const double t_ps = (double)(st + ps) * PERIOD, t_pe = (double)(st + pe) * PERIOD;
    const double t_peak = 0.5 * (t_ps + t_pe);
    const double ye = yps * FM_E;
    FitIn f;
    f.y = y;
    f.fs = fs;
    f.fe = fe;
    f.start_tdc = st;
    f.lo[0] = ye * p.min_pulse_height_ratio;
    f.lo[1] = t_ps - p.max_time_difference;
    f.lo[2] = p.min_beta;
    f.hi[0] = ye * p.max_pulse_height_ratio;
    f.hi[1] = t_pe + p.max_time_difference;
    f.hi[2] = p.max_beta;
    const double p0[3] = {ye, t_peak, p.default_beta};
    FitOut fo;
    fit_gumbel(p, f, p0, fo);
    const int ndf = (fe - fs + 1) - 3;
    bool near = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double span = f.hi[i] - f.lo[i];
      near |= (fo.par[i] - f.lo[i]) / span < p.limit_tolerance || (f.hi[i] - fo.par[i]) / span < p.limit_tolerance;
    }
    const bool failed = !fo.conv || !fo.hess_ok || near;
    const double A = fo.par[0], mu = fo.par[1], bb = fo.par[2];
    const double pe_ = (failed ? ye * p.default_beta : A * bb) / area[w];
    const double tpulse = failed ? t_peak : mu;
    const double hgt = failed ? yps : A / FM_E;
    const double beta = failed ? p.default_beta : bb;
    const double c2 = (failed || ndf <= 0) ? -1.0 : fo.chi2 / (double)(ndf > 1 ? ndf : 1);
    const bool fseq = failed || failed_before;
    if (!drop) {
      if (np < pcap) {
        int64_t q = pb + np;
        o.seq[q] = seq;
        o.pe[q] = pe_;
        o.peph[q] = hgt / height[w];
        o.time[q] = p.use_time_offset_db ? tpulse + toff[w] * p.time_offset_scale : tpulse;
        o.height[q] = hgt;
        o.beta[q] = beta;
        o.chi2[q] = c2;
        o.letime[q] = tpulse - p.le_time_factor * beta;
        o.failed[q] = fseq;
        o.zerondf[q] = ndf <= 0;
        o.ped[q] = ped;
        o.pedfromdb[q] = !und;
        o.fa[q] = A;
        o.fmu[q] = mu;
        o.fb[q] = bb;
      }
      ++np;
    }
    failed_before = fseq;
    for (int i = 0; i < n; ++i) {
      const double tt = (double)(st + i) * PERIOD;
      const double z = (tt - mu) / bb;
      y[i] = y[i] - round_half_away(A * fm_exp(-z - fm_exp(-z)));
    }
    startb = fe;
  }
  npul[w] = np;
}

__global__ void k_pulse_caps(int64_t nwf, const int32_t *__restrict__ nsamp, int64_t *__restrict__ cap) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nwf) cap[w] = nsamp[w] / 2 + 1;
}

__global__ void k_pulses_out(int64_t nwf, const int64_t *__restrict__ ev, const int32_t *__restrict__ chan,
                             const int64_t *__restrict__ poff, const int64_t *__restrict__ pout, PulseSlots s,
                             crv_pulses o) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwf) return;
  int64_t src = poff[w], dst = pout[w], n = pout[w + 1] - dst;
  for (int64_t k = 0; k < n; ++k) {
    int64_t a = src + k, b = dst + k;
    o.wf[b] = w;
    o.seq[b] = s.seq[a];
    if (o.ev) o.ev[b] = ev ? ev[w] : 0;
    if (o.chan) o.chan[b] = chan ? chan[w] : 0;
    o.pe[b] = s.pe[a];
    o.peph[b] = s.peph[a];
    o.time[b] = s.time[a];
    o.height[b] = s.height[a];
    o.beta[b] = s.beta[a];
    o.chi2[b] = s.chi2[a];
    o.letime[b] = s.letime[a];
    o.ped[b] = s.ped[a];
    o.fit_a[b] = s.fa[a];
    o.fit_mu[b] = s.fmu[a];
    o.fit_b[b] = s.fb[a];
    o.failed[b] = s.failed[a];
    o.zerondf[b] = s.zerondf[a];
    o.pedfromdb[b] = s.pedfromdb[a];
  }
}

}  // namespace

namespace {

// Workspace of the allocating entry points: the context's slots (grown on demand).
struct WsAlloc {
  Ctx *c;
  void *operator()(int slot, size_t bytes) { return ws(c, slot, bytes); }
  int check(int64_t, int64_t) { return CRV_OK; }
  int fail() { return ws_fail(c); }
  int scan64(const int64_t *in, int64_t *out, int64_t n, int slot, hipStream_t s) {
    return scan_i64(c, in, out, n, slot, s);
  }
  int scan32(const int32_t *in, int64_t *out, int64_t n, int slot, hipStream_t s) {
    return scan_i32_to_i64(c, in, out, n, slot, s);
  }
  static constexpr bool internal_outputs = true;
};

// ---- scratch of the allocation-free entry point (crv_pulses_run_noalloc) ----
// One fixed device array per GPU, carved in the same order by every call; the sizes below bound one call.
constexpr int64_t NA_MAX_WF = 1 << 17;       // reco waveforms per call
constexpr int64_t NA_MAX_SAMPLES = 1 << 19;  // samples per call (all waveforms together)
constexpr int64_t NA_MAX_SLOTS = NA_MAX_SAMPLES / 2 + NA_MAX_WF;  // pulse slots: nsamp / 2 + 1 per waveform
constexpr size_t NA_SCAN_TMP = 256 << 10;    // bytes for the scan temporaries (checked at set-up)
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t NA_BYTES = al256((NA_MAX_SAMPLES + 1) * 8)          // pedestal-subtracted samples
                            + 4 * al256((NA_MAX_WF + 1) * 8)          // slot counts, slot offsets, pulse offsets, widened
                            + al256((NA_MAX_WF + 1) * 4)              // pulses per waveform
                            + 12 * al256((NA_MAX_SLOTS + 1) * 8)      // slot seq + 11 FP64 columns
                            + 3 * al256(NA_MAX_SLOTS + 8)             // slot flags
                            + NA_SCAN_TMP;

static __device__ __align__(256) uint8_t scratch[NA_BYTES];

std::mutex g_scratch_mu;
std::map<int, uint8_t *> g_scratch;  // device -> address of that device's copy of `scratch`

__global__ void k_zero_first(int64_t *__restrict__ out) { *out = 0; }

__global__ void k_widen32(const int32_t *__restrict__ in, int64_t *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

struct FixedAlloc {
  Ctx *c;
  uint8_t *base;
  size_t used = 0;
  void *operator()(int, size_t bytes) {
    const size_t b = al256(bytes ? bytes : 8);
    if (used + b > NA_BYTES) return nullptr;
    void *p = base + used;
    used += b;
    return p;
  }
  int check(int64_t n_wf, int64_t n_samples) {
    if (n_wf > NA_MAX_WF || n_samples > NA_MAX_SAMPLES)
      return set_error(c, CRV_E_ARG, "pulses (no-alloc): " + std::to_string(n_wf) + " waveforms / " +
                                         std::to_string(n_samples) + " samples exceed the scratch limit of " +
                                         std::to_string(NA_MAX_WF) + " / " + std::to_string(NA_MAX_SAMPLES));
    return CRV_OK;
  }
  int fail() { return set_error(c, CRV_E_STATE, "pulses (no-alloc): scratch exhausted"); }
  int scan64(const int64_t *in, int64_t *out, int64_t n, int, hipStream_t s) {
    k_zero_first<<<1, 1, 0, s>>>(out);
    CRV_HIP(c, hipGetLastError());
    if (n == 0) return CRV_OK;
    size_t tb = 0;
    CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(nullptr, tb, in, out + 1, (int)n, s));
    void *tmp = (*this)(0, tb);
    if (!tmp || tb > NA_SCAN_TMP) return fail();
    CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(tmp, tb, in, out + 1, (int)n, s));
    return CRV_OK;
  }
  int scan32(const int32_t *in, int64_t *out, int64_t n, int slot, hipStream_t s) {
    int64_t *w = (int64_t *)(*this)(0, (n + 1) * 8);
    if (!w) return fail();
    if (n) k_widen32<<<nblocks(n, 256), 256, 0, s>>>(in, w, n);
    CRV_HIP(c, hipGetLastError());
    return scan64(w, out, n, slot, s);
  }
  static constexpr bool internal_outputs = false;
};

template <class Alloc>
int pulses_impl(Ctx *c, Alloc &A, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                const double *height, const double *toff, crv_pulses *out, hipStream_t s) {
  if (n_wf == 0) {
    out->n = 0;
    return CRV_OK;
  }
  int rc = A.check(n_wf, 0);
  if (rc) return rc;
  int64_t nsamp_total = 0;
  if ((rc = read_i64(c, off + n_wf, &nsamp_total, s))) return rc;
  if ((rc = A.check(n_wf, nsamp_total))) return rc;
  double *yscr = (double *)A(60, (nsamp_total + 1) * 8);
  int64_t *cap = (int64_t *)A(61, (n_wf + 1) * 8);
  int64_t *poff = (int64_t *)A(62, (n_wf + 1) * 8);
  int32_t *npul = (int32_t *)A(63, (n_wf + 1) * 4);
  int64_t *pout = (int64_t *)A(64, (n_wf + 1) * 8);
  if (!yscr || !cap || !poff || !npul || !pout) return A.fail();
  k_pulse_caps<<<nblocks(n_wf, 256), 256, 0, s>>>(n_wf, nsamp, cap);
  CRV_HIP(c, hipGetLastError());
  if ((rc = A.scan64(cap, poff, n_wf, 65, s))) return rc;
  int64_t ncap = 0;
  if ((rc = read_i64(c, poff + n_wf, &ncap, s))) return rc;
  PulseSlots sl;
  sl.seq = (int64_t *)A(66, (ncap + 1) * 8);
  double **dd[] = {&sl.pe, &sl.peph, &sl.time, &sl.height, &sl.beta, &sl.chi2, &sl.letime, &sl.ped, &sl.fa, &sl.fmu, &sl.fb};
  for (int i = 0; i < 11; ++i) {
    *dd[i] = (double *)A(67 + i, (ncap + 1) * 8);
    if (!*dd[i]) return A.fail();
  }
  sl.failed = (uint8_t *)A(78, ncap + 8);
  sl.zerondf = (uint8_t *)A(79, ncap + 8);
  sl.pedfromdb = (uint8_t *)A(80, ncap + 8);
  if (!sl.seq || !sl.failed || !sl.zerondf || !sl.pedfromdb) return A.fail();
  kmark(c, 2, 0, s);
  k_pulses<<<nblocks(n_wf, 64), 64, 0, s>>>(c->p, c->t, n_wf, chan, start, nsamp, off, adc, ped, area, height, toff,
                                            yscr, poff, sl, npul);
  CRV_HIP(c, hipGetLastError());
  kmark(c, 2, 1, s);
  if ((rc = A.scan32(npul, pout, n_wf, 81, s))) return rc;
  int64_t np = 0;
  if ((rc = read_i64(c, pout + n_wf, &np, s))) return rc;
  out->n = np;
  if (out->capacity < 0) {  // internal outputs of the chained run
    if (!Alloc::internal_outputs) return set_error(c, CRV_E_ARG, "pulses (no-alloc): the caller passes the outputs");
    out->wf = (int64_t *)ws(c, 320, np * 8 + 8);
    out->seq = (int64_t *)ws(c, 321, np * 8 + 8);
    out->ev = (int64_t *)ws(c, 322, np * 8 + 8);
    out->chan = (int32_t *)ws(c, 323, np * 4 + 8);
    double **dd2[] = {&out->pe, &out->peph, &out->time, &out->height, &out->beta, &out->chi2, &out->letime, &out->ped,
                      &out->fit_a, &out->fit_mu, &out->fit_b};
    for (int i = 0; i < 11; ++i) *dd2[i] = (double *)ws(c, 324 + i, np * 8 + 8);
    out->failed = (uint8_t *)ws(c, 335, np + 8);
    out->zerondf = (uint8_t *)ws(c, 336, np + 8);
    out->pedfromdb = (uint8_t *)ws(c, 337, np + 8);
    out->capacity = np;
  }
  if (np > out->capacity) return set_error(c, CRV_E_CAPACITY, "pulses: output capacity too small");
  k_pulses_out<<<nblocks(n_wf, 256), 256, 0, s>>>(n_wf, ev, chan, poff, pout, sl, *out);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

}  // namespace

int pulses_stage(Ctx *c, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                 const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                 const double *height, const double *toff, crv_pulses *out, hipStream_t s) {
  WsAlloc A{c};
  return pulses_impl(c, A, ev, chan, start, nsamp, off, adc, n_wf, ped, area, height, toff, out, s);
}

int noalloc_setup(Ctx *c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(c->device);
  if (it == g_scratch.end()) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(c, CRV_E_HIP, std::string("no-alloc scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    size_t tb = 0;  // the scan temporaries of the largest call must fit their share
    CRV_HIP(c, hipcub::DeviceScan::InclusiveSum(nullptr, tb, (const int64_t *)nullptr, (int64_t *)nullptr,
                                               (int)NA_MAX_WF, (hipStream_t)0));
    if (tb > NA_SCAN_TMP) return set_error(c, CRV_E_STATE, "no-alloc scratch: scan temporaries do not fit");
    it = g_scratch.emplace(c->device, (uint8_t *)p).first;
  }
  c->scratch = it->second;
  return CRV_OK;
}

int pulses_stage_noalloc(Ctx *c, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                         const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                         const double *height, const double *toff, crv_pulses *out, hipStream_t s) {
  if (!c->scratch) return set_error(c, CRV_E_STATE, "pulses (no-alloc): scratch not set up on this device");
  FixedAlloc A{c, c->scratch};
  return pulses_impl(c, A, ev, chan, start, nsamp, off, adc, n_wf, ped, area, height, toff, out, s);
}

}  // namespace crv
