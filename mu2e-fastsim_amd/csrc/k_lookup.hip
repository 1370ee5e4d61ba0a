// Photon arrivals at the SiPMs through the v6 lookup tables (crv_lookup_photons_run), FP64.
//
// Follows the reference's response/photons.py:
//  * rows (iter_rows, :305-356): per step and simulated SiPM s its local start and direction in the bar (mirrored:
//    s = 1 z -> -z, 2 y -> -y, 3 both; reflector sectors simulate s = 0 and 2 only, reflector < 0 flips z first
//    and writes SiPM s to channel s + 1, its yield deviation still that of s), the speed-ratio delay k (z + H)
//    added to both endpoint times, the attenuation-ratio slope mu, beta (:80-85) and the three photon numbers
//    (scintillation Y edep, scintillator and fibre Cerenkov per mm x |q| x length, :88-92) drawn Gauss(avg,
//    sqrt avg) above 10 else Poisson, divided by the substep count and rounded (:102-112, :342);
//  * substeps (_substep_photons, :163-224): position and time at the centre of each 0.1 mm substep; inside the
//    scintillator (:139-146) the scintillation and scintillator-Cerenkov bins, else in a fibre (:149-160) the fibre
//    bin; p = prob * dev * e^(-mu z), capped at 1; Binomial(n, p) arrivals, each at the substep time + a delay index
//    drawn from the bin's weights (the first j with u * sum <= cum_j, :115-131);
//  * the stock SiPM-stage mask (chain.py:199-204): channels with status bit 0 or 2 give no photons.
// Draws are keyed by Philox on (event, channel, step ordinal, ...) as the oracle lays them out: photon numbers
// (KIND_LCOUNT, channel + type * 2^15), Binomial chunks per substep (KIND_LBINOM: call 0 holds chunk 0 of the three
// types, chunk c >= 1 of type t is call 1 + 3 ((c - 1) >> 2) + t, word (c - 1) & 3, channel + call * 2^15), delays
// per photon (KIND_LDELAY, rank >> 2, word rank & 3).
//
// Layout: one wavefront per (step, SiPM slot) row, its lanes over the row's substeps 64 at a time. A count pass
// reduces each row's arrivals, a scan places the rows, and the fill pass recomputes the same draws and writes each
// lane's arrivals at its wavefront prefix-scan offset, so a row's photons come out ranked substep by substep, then by
// type (scintillation, scintillator Cerenkov, fibre Cerenkov), then by arrival. Bin edges, probabilities and the
// cumulative delay weights stay in HBM (read through L2; a table is a few hundred MB at most).
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "crv_internal.h"

namespace crv {

namespace {

constexpr uint32_t KIND_LCOUNT = 12, KIND_LBINOM = 13, KIND_LDELAY = 14;
constexpr double PRECISION_MM = 0.1, GAUSS_ABOVE = 10.0, BINOM_CHUNK_MEAN = 16.0;
constexpr int64_t MAX_SUBSTEPS = 1 << 24, BINOM_MAX_CHUNKS = 1 << 16, BINOM_KMAX = 1024;
// bits of the lookup stage's flag word
enum LkBad { LK_BAR = 1, LK_LONG = 2, LK_AVG = 4, LK_CHUNKS = 8 };
constexpr int F_LK = 7;

struct Row {
  bool on;
  uint32_t evlo, chan, j;
  int64_t step, nsub;
  int capped;  // Poisson draws of the row's photon numbers that hit the rejection bound
  const LkTab *T;
  double ax, ay, az, dx, dy, dz, ta, tb, beta, theta, dev, mu;
  int64_t n[3];
};

__device__ __forceinline__ int find_bin(const double *__restrict__ e, int ne, double v) {
  // the first i with e[i] <= v <= e[i + 1]: lower bound of v, moved to 1 on the first edge
  int lo = 0, hi = ne;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  if (v == e[0]) lo = 1;
  return (lo >= 1 && lo <= ne - 1) ? lo - 1 : -1;
}

__device__ __forceinline__ double hyp(double a, double b) { return sqrt(a * a + b * b); }

__device__ __forceinline__ double interp_map(double beta, const double *__restrict__ m, int n) {
  // photons per mm: (0, 0) then the map's points (beta, value), flat beyond the last
  if (beta <= 0.0) return 0.0;
  if (n == 0) return 0.0;
  if (beta >= m[2 * (n - 1)]) return m[2 * (n - 1) + 1];
  double b0 = 0.0, v0 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double b1 = m[2 * i], v1 = m[2 * i + 1];
    if (beta < b1) return v0 + (beta - b0) * ((v1 - v0) / (b1 - b0));
    b0 = b1;
    v0 = v1;
  }
  return v0;
}

__device__ __forceinline__ double fm_acos(double x) {
  const double a = fabs(x);
  if (!(a <= 1.0)) return NAN;
  if (a <= 0.5) return FM_PI_2 - x * horner(Coef::asin_c, x * x);
  const double w = (1.0 - a) * 0.5;
  const double v = 2.0 * (sqrt(w) * horner(Coef::asin_c, w));
  return x > 0 ? v : (2.0 * FM_PI_2) - v;
}

// photons per substep of one type (reference _per_substep): Gauss above 10, else Poisson
__device__ __forceinline__ int64_t per_substep(double avg, int64_t nsub, const Draw &dr, uint32_t ev, uint32_t c1,
                                               uint32_t j, int *flags, int &capped_draws) {
  double x;
  if (!isfinite(avg)) {
    atomicOr(&flags[F_LK], LK_AVG);
    return 0;
  }
  if (avg > GAUSS_ABOVE) {
    const U4 u = dr(KIND_LCOUNT, ev, c1, j, 0);
    x = avg + sqrt(avg) * box_muller(unif(u.x), unif(u.y));
  } else {
    bool capped = false;
    x = (double)poisson(avg > 0.0 ? avg : 0.0, dr, KIND_LCOUNT, ev, c1, j, &capped);
    capped_draws += capped ? 1 : 0;
  }
  const double r = rint(x / (double)nsub);
  return r > 0.0 ? (int64_t)r : 0;
}

__device__ __forceinline__ bool row_setup(Row &R, int64_t row, const crv_lookup_steps &st, const LkBars &B,
                          const LkTab *__restrict__ tabs, const crv_params &p, const Tables &t, int *flags) {
  R.on = false;
  const int64_t i = row >> 2;
  const int s = (int)(row & 3);
  const int bar = st.bar[i];
  if (bar < 0 || bar >= B.nbars) {
    if (s == 0) atomicOr(&flags[F_LK], LK_BAR);
    return false;
  }
  const int tid = B.table[bar];
  if (tid < 0) return false;
  const int refl = B.refl[bar];
  if (refl != 0 && (s & 1)) return false;
  const double t0 = st.t0[i], t1 = st.t1[i];
  if (!isfinite(t0) || !isfinite(t1)) return false;
  const int chan = bar * 4 + (refl < 0 ? s + 1 : s), chdev = bar * 4 + s;
  if (chan >= t.n_channels || chdev >= t.n_channels || chan >= CRV_CHAN_MULT) {
    atomicAdd(&flags[F_BADCHAN], 1);
    return false;
  }
  if (t.status[chan] & 0b101) return false;
  const uint32_t ev = st.event[i];
  if (i > 0 && st.event[i - 1] > ev) atomicAdd(&flags[F_UNSORTED], 1);
  // step ordinal in its event: the lower bound of the event in the sorted rows
  int64_t lo = 0, hi = i;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (st.event[mid] < ev) lo = mid + 1;
    else hi = mid;
  }
  const LkTab *T = tabs + tid;
  // local coordinates (thickness, width, length) of the step's ends in the bar
  double a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    const int w = B.l2w[bar * 3 + k];
    const double c = B.center[bar * 3 + w];
    a[k] = st.pos0[i * 3 + w] - c;
    b[k] = st.pos1[i * 3 + w] - c;
  }
  if (refl < 0) {
    a[2] = -a[2];
    b[2] = -b[2];
  }
  if (s == 1 || s == 3) {
    a[2] = -a[2];
    b[2] = -b[2];
  }
  if (s == 2 || s == 3) {
    a[1] = -a[1];
    b[1] = -b[1];
  }
  const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  const double dl = sqrt((dx * dx + dy * dy) + dz * dz);
  R.theta = dl > 0 ? fm_acos(fmin(fmax(dz / dl, -1.0), 1.0)) : 0.0;
  const double q = dl / PRECISION_MM;
  R.nsub = (q >= 1.0 && q < 9.2e18) ? (int64_t)floor(q) : 1;
  if (R.nsub > MAX_SUBSTEPS) {
    atomicOr(&flags[F_LK], LK_LONG);
    return false;
  }
  const double k = T->inv_v * (1.0 / B.speed[bar] - 1.0);
  R.mu = (1.0 / B.atten[bar] - 1.0) / T->att;
  R.ta = t0 + k * (a[2] + T->hl);
  R.tb = t1 + k * (b[2] + T->hl);
  R.dev = t.dev[chdev];
  const double m = st.mass[i], p0 = st.p0[i], p1 = st.p1[i];
  if (m > 0) {
    const double e = 0.5 * (sqrt(p0 * p0 + m * m) + sqrt(p1 * p1 + m * m));
    const double g = e / m;
    R.beta = sqrt(fmin(fmax(1.0 - 1.0 / (g * g), 0.0), 1.0));
  } else {
    R.beta = 1.0;
  }
  const double qabs = fabs(st.charge[i]), len = st.len[i];
  const Draw dr{(uint32_t)p.seed, p.evkey_hi};
  const uint32_t j = (uint32_t)(i - lo);
  const double avg[3] = {B.yield[bar] * st.edep[i], (interp_map(R.beta, T->ck[0], T->nck[0]) * qabs) * len,
                         (interp_map(R.beta, T->ck[1], T->nck[1]) * qabs) * len};
  R.capped = 0;
#pragma unroll
  for (int ty = 0; ty < 3; ++ty)
    R.n[ty] = per_substep(avg[ty], R.nsub, dr, ev, (uint32_t)chan + ((uint32_t)ty << 15), j, flags, R.capped);
  R.on = true;
  R.evlo = ev;
  R.chan = (uint32_t)chan;
  R.j = j;
  R.step = i;
  R.T = T;
  R.ax = a[0];
  R.ay = a[1];
  R.az = a[2];
  R.dx = dx;
  R.dy = dy;
  R.dz = dz;
  return true;
}

struct Sub {
  double t;
  int32_t bin[3];
  int64_t n[3];
};

// Binomial(n, p) by chunked inversion (oracle.rand.binomial): u0 the call-0 draw of the substep
__device__ int64_t binomial(int64_t n, double p, const U4 &u0, int ty, const Draw &dr, uint32_t ev, uint32_t chan,
                            uint32_t j, uint32_t sub, int *flags) {
  const bool flip = p > 0.5;
  const double q = flip ? 1.0 - p : p;
  if (n <= 0) return 0;
  if (!(q > 0.0)) return flip ? n : 0;
  const double lq = fm_log(1.0 - q);
  const double lim = BINOM_CHUNK_MEAN / -lq;
  int64_t m = lim >= (double)n ? n : (int64_t)fmax(floor(lim), 1.0);
  const int64_t nch = (n + m - 1) / m;
  if (nch > BINOM_MAX_CHUNKS) {
    atomicOr(&flags[F_LK], LK_CHUNKS);
    return 0;
  }
  const double r = q / (1.0 - q);
  int64_t tot = 0;
  U4 cache = u0;
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t size = m < n - c * m ? m : n - c * m;
    uint32_t w;
    if (c == 0) {
      w = ty == 0 ? u0.x : ty == 1 ? u0.y : u0.z;
    } else {
      const int64_t c1 = c - 1;
      if ((c1 & 3) == 0) cache = dr(KIND_LBINOM, ev, chan + ((uint32_t)(1 + 3 * (c1 >> 2) + ty) << 15), j, sub);
      const int lane = (int)(c1 & 3);
      w = lane == 0 ? cache.x : lane == 1 ? cache.y : lane == 2 ? cache.z : cache.w;
    }
    const double u = unif(w);
    double f = fm_exp((double)size * lq);
    double cdf = f;
    int64_t k = 0;
    while (u > cdf && k < size && k < BINOM_KMAX) {
      f = ((f * r) * (double)(size - k)) / (double)(k + 1);
      k += 1;
      cdf = cdf + f;
    }
    tot += k;
  }
  return flip ? n - tot : tot;
}

__device__ __forceinline__ void substep(Sub &S, const Row &R, int64_t k, const Draw &dr, int *flags) {
  const LkTab *T = R.T;
  const double frac = ((double)k + 0.5) / (double)R.nsub;
  const double x = R.ax + R.dx * frac, y = R.ay + R.dy * frac, z = R.az + R.dz * frac;
  S.t = R.ta + (R.tb - R.ta) * frac;
  const double ax = fabs(x), ay = fabs(y);
  bool insc = ax < T->ht && ay < T->hw && fabs(z) < T->hl;
  const double ex = ax / T->hrx, ey = (ay - T->fs / 2) / T->hry;
  insc = insc && (ex * ex + ey * ey) >= 1.0;
  const double cx = T->ht - T->cr, cy = T->hw - T->cr;
  if (ax > cx && ay > cy && hyp(ax - cx, ay - cy) > T->cr) insc = false;
  const double h = T->fs / 2;
  const double r0 = hyp(x, y + h), r1 = hyp(x, y - h);
  const bool in1 = r1 <= T->fr;
  const bool infb = !insc && fabs(z) < T->hl && (r0 <= T->fr || in1);
  const int *nb = T->nb;
  const int bz = find_bin(T->edges[2], T->ne[2], z);
  const int bb = find_bin(T->edges[3], T->ne[3], R.beta);
  S.bin[0] = S.bin[1] = S.bin[2] = -1;
  if (insc) {
    const int bx = find_bin(T->edges[0], T->ne[0], ax), by = find_bin(T->edges[1], T->ne[1], y);
    if (bx >= 0 && by >= 0 && bz >= 0) {
      S.bin[0] = bz + by * nb[2] + bx * nb[1] * nb[2];
      if (bb >= 0) S.bin[1] = bb + bz * nb[3] + by * nb[2] * nb[3] + bx * nb[1] * nb[2] * nb[3];
    }
  } else if (infb) {
    const double ry = in1 ? y - h : y + h, r = in1 ? r1 : r0;
    const double norm = hyp(R.dx, R.dy) * r;
    const double phi = norm > 0 ? fm_acos(fmin(fmax((R.dx * x + R.dy * ry) / norm, -1.0), 1.0)) : 0.0;
    const int bth = find_bin(T->edges[4], T->ne[4], R.theta), bph = find_bin(T->edges[5], T->ne[5], phi),
              br = find_bin(T->edges[6], T->ne[6], r);
    if (bb >= 0 && bth >= 0 && bph >= 0 && br >= 0 && bz >= 0)
      S.bin[2] = bz + br * nb[2] + bph * nb[6] * nb[2] + bth * nb[5] * nb[6] * nb[2] + bb * nb[4] * nb[5] * nb[6] * nb[2];
  }
  S.n[0] = S.n[1] = S.n[2] = 0;
  if ((S.bin[0] < 0 || R.n[0] == 0) && (S.bin[1] < 0 || R.n[1] == 0) && (S.bin[2] < 0 || R.n[2] == 0)) return;
  const U4 u0 = dr(KIND_LBINOM, R.evlo, R.chan, R.j, (uint32_t)k);
  const double att = R.mu != 0.0 ? fm_exp(-R.mu * z) : 1.0;
#pragma unroll
  for (int ty = 0; ty < 3; ++ty) {
    if (S.bin[ty] < 0 || R.n[ty] == 0) continue;
    double pr = (double)T->prob[ty][S.bin[ty]] * R.dev;
    if (R.mu != 0.0) pr = pr * att;
    pr = pr > 1.0 ? 1.0 : pr;
    S.n[ty] = binomial(R.n[ty], pr, u0, ty, dr, R.evlo, R.chan, R.j, (uint32_t)k, flags);
  }
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ void __launch_bounds__(256) k_lk_count(crv_lookup_steps st, LkBars B, const LkTab *__restrict__ tabs,
                                                  crv_params p, Tables t, int64_t *__restrict__ rcount,
                                                  int *__restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int64_t nrows = st.n * 4, nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const Draw dr{(uint32_t)p.seed, p.evkey_hi};
  for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); row < nrows; row += nw) {
    Row R;
    int64_t tot = 0;
    if (row_setup(R, row, st, B, tabs, p, t, flags)) {
      for (int64_t k0 = 0; k0 < R.nsub; k0 += 64) {
        int64_t c = 0;
        if (k0 + lane < R.nsub) {
          Sub S;
          substep(S, R, k0 + lane, dr, flags);
          c = S.n[0] + S.n[1] + S.n[2];
        }
        tot += wave_sum(c);
      }
      if (tot > 0xFFFF) {
        if (lane == 0) atomicAdd(&flags[F_OVERFLOW], 1);
        tot = 0;
      }
      if (lane == 0 && R.capped) atomicAdd(&flags[F_CAPPED], R.capped);
    }
    if (lane == 0) rcount[row] = tot;
  }
}

__global__ void __launch_bounds__(256) k_lk_fill(crv_lookup_steps st, LkBars B, const LkTab *__restrict__ tabs,
                                                 crv_params p, Tables t, const int64_t *__restrict__ roff,
                                                 crv_photons out, int *__restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int64_t nrows = st.n * 4, nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const Draw dr{(uint32_t)p.seed, p.evkey_hi};
  const int64_t evhi = (int64_t)p.evkey_hi << 32;
  for (int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); row < nrows; row += nw) {
    if (roff[row + 1] == roff[row]) continue;  // no photons (also every inactive row)
    Row R;
    if (!row_setup(R, row, st, B, tabs, p, t, flags)) continue;
    const int64_t base = roff[row];
    const int64_t label = (evhi | (int64_t)R.evlo) * CRV_CHAN_MULT + R.chan;
    int64_t rank0 = 0;
    for (int64_t k0 = 0; k0 < R.nsub; k0 += 64) {
      Sub S;
      int64_t c = 0;
      const bool mine = k0 + lane < R.nsub;
      if (mine) {
        substep(S, R, k0 + lane, dr, flags);
        c = S.n[0] + S.n[1] + S.n[2];
      }
      int64_t incl = c;
      for (int o = 1; o < 64; o <<= 1) {
        const int64_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
      }
      int64_t r = rank0 + incl - c;
      rank0 += __shfl(incl, 63);
      if (!mine || c == 0) continue;
      U4 cache{};
      int64_t cached = -1;
#pragma unroll
      for (int ty = 0; ty < 3; ++ty) {
        if (!S.n[ty]) continue;
        const int dmax = R.T->dmax[ty];
        const int32_t *cum = R.T->cum[ty] + (int64_t)S.bin[ty] * dmax;
        const double sum = (double)R.T->dsum[ty][S.bin[ty]];
        for (int64_t a = 0; a < S.n[ty]; ++a, ++r) {
          if ((r >> 2) != cached) {
            cached = r >> 2;
            cache = dr(KIND_LDELAY, R.evlo, R.chan, R.j, (uint32_t)cached);
          }
          const int w = (int)(r & 3);
          const double thr = unif(w == 0 ? cache.x : w == 1 ? cache.y : w == 2 ? cache.z : cache.w) * sum;
          int lo = 0, hi = dmax;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((double)cum[mid] < thr) lo = mid + 1;
            else hi = mid;
          }
          const int64_t o = base + r;
          out.label[o] = label;
          out.time[o] = S.t + (double)lo;
          out.photon[o] = R.step;
          out.j[o] = R.j;
          out.rank[o] = (uint32_t)r;
        }
      }
    }
  }
}

__global__ void k_lk_cum(const uint8_t *__restrict__ w, int64_t nbins, int dmax, int32_t *__restrict__ cum) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int32_t s = 0;
  for (int i = 0; i < dmax; ++i) {
    s += w[b * dmax + i];
    cum[b * dmax + i] = s;
  }
}

unsigned lk_grid(int64_t rows) {
  const int64_t waves = std::min<int64_t>(rows, 256 * 32);
  return (unsigned)std::max<int64_t>((waves + 3) / 4, 1);
}

}  // namespace

void lookup_free(Ctx *c) {
  for (void *q : c->lk_owned) (void)hipFree(q);
  c->lk_owned.clear();
  c->lk_tabs = nullptr;
  c->lk_ntab = 0;
  c->lk_bars = LkBars{};
}

int lookup_upload(Ctx *c, const crv_lookup_table *tabs, int32_t ntab, const crv_lookup_bars *bars) {
  if (ntab < 1 || ntab > 16 || !tabs || !bars) return set_error(c, CRV_E_ARG, "lookup: 1-16 tables and the bars");
  if (bars->nbars <= 0 || !bars->table || !bars->reflector || !bars->yield || !bars->speed_ratio ||
      !bars->atten_ratio || !bars->center || !bars->local_to_world)
    return set_error(c, CRV_E_ARG, "lookup: incomplete bar arrays");
  lookup_free(c);
  auto up = [&](const void *h, size_t bytes) -> void * {
    void *d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(bytes, 8)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    c->lk_owned.push_back(d);
    if (bytes && h && hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  };
  std::vector<LkTab> host(ntab);
  for (int i = 0; i < ntab; ++i) {
    const crv_lookup_table &L = tabs[i];
    LkTab &T = host[i];
    T.ht = L.half_thickness;
    T.hw = L.half_width;
    T.hl = L.half_length;
    T.fs = L.fiber_separation;
    T.hrx = L.hole_radius_x;
    T.hry = L.hole_radius_y;
    T.fr = L.fiber_radius;
    T.cr = L.corner_radius;
    T.inv_v = L.inv_speed;
    T.att = L.atten_length;
    for (int k = 0; k < 7; ++k) {
      if (L.n_edges[k] < 2 || !L.edges[k]) return set_error(c, CRV_E_ARG, "lookup: every axis needs two edges");
      T.ne[k] = L.n_edges[k];
      T.nb[k] = L.n_edges[k] - 1;
      if (!(T.edges[k] = (const double *)up(L.edges[k], (size_t)L.n_edges[k] * 8))) goto fail;
    }
    for (int k = 0; k < 2; ++k) {
      if (L.n_ck[k] < 0 || (L.n_ck[k] && !L.ck[k])) return set_error(c, CRV_E_ARG, "lookup: Cerenkov map");
      T.nck[k] = L.n_ck[k];
      if (!(T.ck[k] = (const double *)up(L.ck[k], (size_t)L.n_ck[k] * 16))) goto fail;
    }
    {
      const int64_t ns = (int64_t)T.nb[0] * T.nb[1] * T.nb[2];
      const int64_t nbin[3] = {ns, ns * T.nb[3], (int64_t)T.nb[3] * T.nb[4] * T.nb[5] * T.nb[6] * T.nb[2]};
      for (int k = 0; k < 3; ++k) {
        if (L.dmax[k] < 1 || !L.prob[k] || !L.delays[k])
          return set_error(c, CRV_E_ARG, "lookup: every table needs probabilities and delay weights");
        if ((int64_t)L.dmax[k] * 255 > INT32_MAX) return set_error(c, CRV_E_ARG, "lookup: delay vector too long");
        T.dmax[k] = L.dmax[k];
        if (!(T.prob[k] = (const float *)up(L.prob[k], (size_t)nbin[k] * 4))) goto fail;
        const uint8_t *dw = (const uint8_t *)up(L.delays[k], (size_t)nbin[k] * L.dmax[k]);
        int32_t *cum = (int32_t *)up(nullptr, (size_t)nbin[k] * L.dmax[k] * 4);
        if (!dw || !cum) goto fail;
        if (nbin[k]) k_lk_cum<<<nblocks(nbin[k], 256), 256>>>(dw, nbin[k], L.dmax[k], cum);
        if (hipGetLastError() != hipSuccess) goto fail;
        // row sums: the last cumulative entry of each bin
        std::vector<int32_t> sums(nbin[k]);
        for (int64_t b = 0; b < nbin[k]; ++b) {
          int32_t s = 0;
          for (int q = 0; q < L.dmax[k]; ++q) s += L.delays[k][b * L.dmax[k] + q];
          sums[b] = s;
        }
        T.cum[k] = cum;
        if (!(T.dsum[k] = (const int32_t *)up(sums.data(), sums.size() * 4))) goto fail;
      }
    }
  }
  {
    const int nb = bars->nbars;
    for (int b = 0; b < nb; ++b) {
      const int tid = bars->table[b];
      if (tid >= ntab) return set_error(c, CRV_E_ARG, "lookup: bar table index beyond the tables");
      for (int k = 0; k < 3; ++k)
        if (bars->local_to_world[b * 3 + k] < 0 || bars->local_to_world[b * 3 + k] > 2)
          return set_error(c, CRV_E_ARG, "lookup: local_to_world entries must be 0-2");
    }
    LkBars &B = c->lk_bars;
    B.nbars = nb;
    B.table = (const int32_t *)up(bars->table, (size_t)nb * 4);
    B.refl = (const int32_t *)up(bars->reflector, (size_t)nb * 4);
    B.yield = (const double *)up(bars->yield, (size_t)nb * 8);
    B.speed = (const double *)up(bars->speed_ratio, (size_t)nb * 8);
    B.atten = (const double *)up(bars->atten_ratio, (size_t)nb * 8);
    B.center = (const double *)up(bars->center, (size_t)nb * 24);
    B.l2w = (const int32_t *)up(bars->local_to_world, (size_t)nb * 12);
    if (!B.table || !B.refl || !B.yield || !B.speed || !B.atten || !B.center || !B.l2w) goto fail;
    c->lk_tabs = (const LkTab *)up(host.data(), host.size() * sizeof(LkTab));
    if (!c->lk_tabs) goto fail;
    c->lk_ntab = ntab;
  }
  if (hipDeviceSynchronize() != hipSuccess) goto fail;
  return CRV_OK;
fail:
  (void)hipGetLastError();
  lookup_free(c);
  return set_error(c, CRV_E_HIP, "lookup: table upload failed");
}

int lookup_photons_stage(Ctx *c, const crv_lookup_steps *st, crv_photons *out, hipStream_t s) {
  out->capped = 0;
  if (!c->lk_tabs) return set_error(c, CRV_E_STATE, "lookup: tables not uploaded");
  if (st->n < 0 || st->n > (INT64_MAX >> 3)) return set_error(c, CRV_E_ARG, "lookup: step count");
  if (st->n == 0) {
    out->n = 0;
    return CRV_OK;
  }
  if (!st->event || !st->bar || !st->t0 || !st->t1 || !st->pos0 || !st->pos1 || !st->edep || !st->len || !st->p0 ||
      !st->p1 || !st->mass || !st->charge)
    return set_error(c, CRV_E_ARG, "lookup: NULL step column");
  const int64_t nrows = st->n * 4;
  int64_t *rcount = (int64_t *)ws(c, 22, (nrows + 1) * 8);
  int64_t *roff = (int64_t *)ws(c, 23, (nrows + 1) * 8);
  if (!rcount || !roff) return ws_fail(c);
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  const unsigned grid = lk_grid(nrows);
  k_lk_count<<<grid, 256, 0, s>>>(*st, c->lk_bars, c->lk_tabs, c->p, c->t, rcount, c->flags);
  CRV_HIP(c, hipGetLastError());
  int rc = scan_i64(c, rcount, roff, nrows, 24, s);
  if (rc) return rc;
  int64_t total = 0;
  if ((rc = read_i64(c, roff + nrows, &total, s))) return rc;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_LK] & LK_BAR) return set_error(c, CRV_E_ARG, "lookup: step bar outside the uploaded bars");
  if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "lookup: channel beyond the uploaded conditions");
  if (hf[F_UNSORTED]) return set_error(c, CRV_E_ARG, "lookup: steps must be sorted by event");
  if (hf[F_LK] & LK_LONG) return set_error(c, CRV_E_ARG, "lookup: a step longer than 2^24 substeps of 0.1 mm");
  if (hf[F_LK] & LK_AVG) return set_error(c, CRV_E_ARG, "lookup: non-finite photon-number average (edep, len, beta)");
  if (hf[F_LK] & LK_CHUNKS)
    return set_error(c, CRV_E_OVERFLOW, "lookup: a substep's Binomial draw needs more chunks than the draw layout");
  if (hf[F_OVERFLOW]) return set_error(c, CRV_E_OVERFLOW, "lookup: more than 65535 photons from one (step, SiPM)");
  out->capped = hf[F_CAPPED];
  out->n = total;
  if (total > out->capacity) return set_error(c, CRV_E_CAPACITY, "lookup: output capacity too small");
  if (total == 0) return CRV_OK;
  if (!out->label || !out->time || !out->photon || !out->j || !out->rank)
    return set_error(c, CRV_E_ARG, "lookup: NULL output column");
  k_lk_fill<<<grid, 256, 0, s>>>(*st, c->lk_bars, c->lk_tabs, c->p, c->t, roff, *out, c->flags);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}

}  // namespace crv
