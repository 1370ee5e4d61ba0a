// Coincidence finder: pulses -> coincidence clusters (FP64): one wavefront per event of up to 256 pulses (the
// event's hits in LDS), one 256-thread block per larger event (scratch in HBM, bucketed and binned searches).
//
// Follows the reference's reco/coinc.py:112-392 and reco/coinc_cpu.py:11-51 step by step: hits from first pulses
// with the fibre-length time correction and partner-status PE compensation; filterHits per readout side (same and
// adjacent counter within max_dt_adjacent); greedy findClusters (repeated passes in list order, a hit joins through
// any member within max(maxd) in x and max_dt in time); checkCoincidence over layer combinations (time window,
// slopes, slope differences, maxDistance widened to the x gaps of passing combinations, kept hits in pulse index
// order); findClusters again across sides; opposite-side attachment (nearest cluster per candidate, one pulse per
// SiPM in (cluster, bar, sipm, dt, candidate) order); cluster moments as sequential sums in member order.
// Outputs go to event-local slots (the event's first pulse onwards), compacted by k_coinc_out.
#include <hipcub/hipcub.hpp>

#include "crv_internal.h"

namespace crv {

namespace {

struct Cut {
  double pe_threshold, max_dt_adjacent, max_dt, min_slope, max_slope, max_slope_diff, coincidence_layers,
      min_cluster_pes, initial_max_distance, final_max_distance;
};

__device__ __forceinline__ Cut cut_of(const Tables &t, int sector) {
  const double *c = t.cuts + (int64_t)sector * 10;
  return Cut{c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9]};
}

struct Ev {  // what the finder needs per pulse: bar, sipm, seq, idx and the derived hit fields
  const int32_t *bar, *sipm;
  const int64_t *seq;
};

struct ClSlots {  // per-cluster outputs of the finder kernels (event-local slots)
  int64_t *cl_ev;
  int32_t *cl_type, *cl_poscalc, *cl_sidehits;
  double *cl_d;  // [slot * 12]
};

// min-PE test and moments of a cluster with members mem[0..mn) (reference reco/coinc.py:291-336 and
// coinc_cpu.py:52-79; every sum runs in member order). A gives pe, corrected time, bar and SiPM of a pulse. Writes the
// cluster to `slot` and returns true, or returns false when the cluster has fewer PE than its sectors' minimum.
template <class Acc>
__device__ bool cluster_emit(const crv_params &p, const Tables &t, const Acc &A, const int32_t *mem, int mn, int type,
                             double P0, double P1, int64_t evkey, int64_t slot, const ClSlots &o) {
  // one pass over the members; every sum still runs in member order
  double tot = 0.0, mincp = INFINITY, acc[3] = {0.0, 0.0, 0.0};
  double sxy = 0.0, sx = 0.0, sy = 0.0, syy = 0.0, cmin = INFINITY, cmax = -INFINITY;
  double ps[2] = {0.0, 0.0}, pt[2] = {0.0, 0.0}, tmn[2] = {0.0, 0.0}, tmx[2] = {0.0, 0.0};
  int nhs[2] = {0, 0};
  for (int i = 0; i < mn; ++i) {
    const int h = mem[i], bar = A.bar(h);
    const double pe = A.pe(h), tt = A.tm(h), x = A.x(h), y = A.y(h);
    tot = tot + pe;
    mincp = fmin(mincp, A.mincp(t, h));
    for (int ax = 0; ax < 3; ++ax) acc[ax] = acc[ax] + A.bp(t, h, ax) * pe;
    sxy = sxy + pe * x * y;
    sx = sx + pe * x;
    sy = sy + pe * y;
    syy = syy + pe * y * y;
    cmin = fmin(cmin, tt);
    cmax = fmax(cmax, tt);
    const int sd = A.sipm(h) % 2;
    if (pe > 0) {
      if (nhs[sd] == 0) {
        tmn[sd] = tt;
        tmx[sd] = tt;
      } else {
        tmn[sd] = fmin(tmn[sd], tt);
        tmx[sd] = fmax(tmx[sd], tt);
      }
      ++nhs[sd];
      ps[sd] = ps[sd] + pe;
      pt[sd] = pt[sd] + tt * pe;
    }
  }
  if (tot < mincp) return false;
  double pos[3], st[2], spe[2];
  for (int ax = 0; ax < 3; ++ax) pos[ax] = acc[ax] / tot;
  for (int sd = 0; sd < 2; ++sd) {
    spe[sd] = ps[sd];
    st[sd] = ps[sd] > 0 ? pt[sd] / ps[sd] : 0.0;
  }
  double slope = (tot * sxy - sx * sy) / (tot * syy - sy * sy);
  int ld = t.ldir[A.sec(mem[0])];
  bool calc = false;
  double avg_t, org;
  const double v = p.fiber_signal_speed;
  if (nhs[0] && nhs[1]) {
    org = 0.5 * (v * (st[0] - st[1]) + (P0 + P1));
    if (P0 < org && org < P1 && tmx[0] - tmn[0] < p.max_time_interval && tmx[1] - tmn[1] < p.max_time_interval) {
      calc = true;
      pos[ld] = org;
      avg_t = st[0] - (org - P0) / v;
    } else {
      org = pos[ld];
      avg_t = 0.5 * ((st[0] - (org - P0) / v) + (st[1] - (P1 - org) / v));
    }
  } else {
    org = pos[ld];
    avg_t = nhs[0] ? (st[0] - (org - P0) / v) : (st[1] - (P1 - org) / v);
  }
  o.cl_ev[slot] = evkey;
  o.cl_type[slot] = type;
  o.cl_poscalc[slot] = calc;
  o.cl_sidehits[slot * 2] = nhs[0];
  o.cl_sidehits[slot * 2 + 1] = nhs[1];
  double *d = o.cl_d + slot * 12;
  d[0] = tot;
  d[1] = cmin;
  d[2] = cmax;
  d[3] = slope;
  d[4] = avg_t - p.time_offset;
  d[5] = pos[0];
  d[6] = pos[1];
  d[7] = pos[2];
  d[8] = spe[0];
  d[9] = spe[1];
  d[10] = st[0];
  d[11] = st[1];
  return true;
}

// ---- wavefront finder: one 64-lane wavefront per event of at most WMAX pulses, the event's hits in LDS ----
// Same steps and arithmetic as k_coinc; the pair tests (filterHits, cluster links, layer combinations, attachment
// candidates) run across the lanes, the order-dependent parts (greedy cluster growth, kept-hit order, attachment
// order, moments in member order) run as in the serial kernel.
template <int WMAX>
struct WS {
  static constexpr int WNW = WMAX / 64;
  double t[WMAX], pe[WMAX], x[WMAX], y[WMAX], maxd[WMAX];
  double wm[WMAX];     // second-clustering max distance, by pulse
  double cmaxd[WMAX];  // widened max distance, by coinc position
  double wq[WMAX];     // widened max distance, by by-layer position
  double cdt[WMAX];    // attachment dt, by candidate position
  uint64_t kmask[WNW], mmask[WNW];
  int32_t bar[WMAX];
  int16_t sec[WMAX], cit[WMAX], chk[WMAX], off[WMAX + 1], off2[WMAX + 1];
  int16_t sk[WMAX];    // sorted attachments: cluster
  uint8_t sc[WMAX];    // sorted attachments: candidate position
  uint8_t layer[WMAX], side[WMAX], sipm[WMAX], seq0[WMAX];
  uint8_t hl[WMAX], filt[WMAX], items[WMAX], coinc[WMAX], byl[WMAX], kept[WMAX], cl2[WMAX];
  uint8_t sl[WMAX], fpass[WMAX];  // hits grouped by (side, layer) in hit order; filterHits result by hit position
  int16_t gof[9];
  int32_t cl_mc[WMAX];  // members of a second-level cluster (0: dropped)
  int32_t mlist[WMAX];  // members of the kept clusters, at their event-local member offsets
  int32_t spre[12];     // layer subsets of a coincidence check: combination prefix counts,
  int16_t slay[11];     // their layers (4 bits each)
  int8_t snl[11];       // and sizes
};

// (bar position and the sector's min_cluster_pes by pulse: staged in LDS arrays that are free by then)
template <int WMAX>
struct WaveAcc {
  const WS<WMAX> &S;
  __device__ double bp(const Tables &, int h, int ax) const { return ax == 0 ? S.wq[h] : ax == 1 ? S.cmaxd[h] : S.wm[h]; }
  __device__ double mincp(const Tables &, int h) const { return S.cdt[h]; }
  __device__ double pe(int h) const { return S.pe[h]; }
  __device__ double tm(int h) const { return S.t[h]; }
  __device__ int bar(int h) const { return S.bar[h]; }
  __device__ int sipm(int h) const { return S.sipm[h]; }
  __device__ int sec(int h) const { return S.sec[h]; }
  __device__ double x(int h) const { return S.x[h]; }
  __device__ double y(int h) const { return S.y[h]; }
};

__device__ __forceinline__ uint64_t lanes_below() { return (1ull << threadIdx.x) - 1ull; }

__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// greedy findClusters of the pulses list[0..m) (reco/coinc.py:128-152): each pass visits the remaining hits in list
// order and a hit joins through any member, including one that joined earlier in the same pass; passes repeat while
// the cluster grows, and the first remaining hit seeds the next cluster. One pass is the least fixpoint of
//   J = { p remaining : link(p) meets M or { q in J : q < p } }      (M = the members before the pass),
// reached from J = {} by repeated evaluation: each lane holds the link rows of its list positions and a round is a
// ballot. Members go to out[] in join order (pass by pass, ascending within a pass), cluster starts to off[].
template <int WMAX>
__device__ int wave_clusters(const Tables &t, WS<WMAX> &S, const uint8_t *list, int m, const double *md, double max_dt,
                             bool opposite, double dt_opp, uint8_t *out, int16_t *off) {
  constexpr int WNW = WMAX / 64;
  const int lane = threadIdx.x;
  const uint64_t below = lanes_below();
  uint64_t row[WNW][WNW];  // link rows of list positions w * 64 + lane
#pragma unroll
  for (int w = 0; w < WNW; ++w) {
#pragma unroll
    for (int v = 0; v < WNW; ++v) row[w][v] = 0ull;
    const int i = w * 64 + lane;
    if (i < m) {
      const int h = list[i];
      const double th = S.t[h], xh = S.x[h], mh = md[h];
      const int sh = S.side[h];
#pragma unroll 4  // the LDS reads of later positions overlap the tests of earlier ones
      for (int j = 0; j < m; ++j) {
        const int c = list[j];
        const double mdd = fmax(mh, md[c]);
        const double mt = max_dt + ((opposite && sh != S.side[c]) ? dt_opp : 0.0);
        const bool l = fabs(xh - S.x[c]) <= mdd && fabs(th - S.t[c]) < mt;
#pragma unroll
        for (int v = 0; v < WNW; ++v)
          if (v == (j >> 6)) row[w][v] |= l ? 1ull << (j & 63) : 0ull;
      }
    }
  }
  uint64_t R[WNW];
#pragma unroll
  for (int w = 0; w < WNW; ++w) {
    const int k = m - w * 64;
    R[w] = k >= 64 ? ~0ull : (k > 0 ? (1ull << k) - 1ull : 0ull);
  }
  int nout = 0, ncl = 0;
  for (;;) {
    int seed = -1;
#pragma unroll
    for (int w = WNW - 1; w >= 0; --w)
      if (R[w]) seed = w * 64 + __ffsll((unsigned long long)R[w]) - 1;
    if (seed < 0) break;
    uint64_t M[WNW];
#pragma unroll
    for (int w = 0; w < WNW; ++w) {
      M[w] = w == (seed >> 6) ? 1ull << (seed & 63) : 0ull;
      R[w] &= ~M[w];
    }
    if (lane == 0) {
      off[ncl] = (int16_t)nout;
      out[nout] = list[seed];
    }
    ++nout;
    for (;;) {  // passes
      uint64_t J[WNW];
#pragma unroll
      for (int w = 0; w < WNW; ++w) J[w] = 0ull;
      for (;;) {  // fixpoint of one pass
        uint64_t Jn[WNW];
        bool same = true;
#pragma unroll
        for (int w = 0; w < WNW; ++w) {
          bool j = false;
          if ((R[w] >> lane) & 1ull) {
#pragma unroll
            for (int v = 0; v < WNW; ++v) {
              const uint64_t lower = v < w ? J[v] : (v == w ? J[v] & below : 0ull);
              j |= (row[w][v] & (M[v] | lower)) != 0ull;
            }
          }
          Jn[w] = __ballot(j);
          same &= Jn[w] == J[w];
        }
#pragma unroll
        for (int w = 0; w < WNW; ++w) J[w] = Jn[w];
        if (same) break;
      }
      bool any = false;
#pragma unroll
      for (int w = 0; w < WNW; ++w) {
        if ((J[w] >> lane) & 1ull) out[nout + __popcll(J[w] & below)] = list[w * 64 + lane];
        nout += __popcll(J[w]);
        M[w] |= J[w];
        R[w] &= ~J[w];
        any |= J[w] != 0ull;
      }
      if (!any) break;
    }
    ++ncl;
  }
  if (lane == 0) off[ncl] = (int16_t)nout;
  __syncthreads();
  return ncl;
}

// the layer combinations of a by-layer list, every needed layer subset in one index space over the lanes (a cluster
// of one track has a handful per subset); UNI: every hit has the cuts `cu`
template <int WMAX, bool UNI>
__device__ void wave_combos(const Tables &t, WS<WMAX> &S, const int lay_off[5], int nmin, int nmax, const Cut &cu) {
  const int lane = threadIdx.x;
  int nsub = 0, total = 0;
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int L[4];
    for (int i = 0; i < nl; ++i) L[i] = i;
    for (;;) {
      int P = 1;
      for (int i = 0; i < nl; ++i) P *= lay_off[L[i] + 1] - lay_off[L[i]];
      if (P > 0) {
        if (lane == 0) {
          S.spre[nsub] = total;
          S.snl[nsub] = (int8_t)nl;
          S.slay[nsub] = (int16_t)(L[0] | L[1] << 4 | (nl > 2 ? L[2] : 0) << 8 | (nl > 3 ? L[3] : 0) << 12);
        }
        ++nsub;
        total += P;
      }
      int k = nl - 1;
      while (k >= 0 && L[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++L[k];
      for (int i = k + 1; i < nl; ++i) L[i] = L[i - 1] + 1;
    }
  }
  if (lane == 0) S.spre[nsub] = total;
  __syncthreads();
  for (int cidx = lane; cidx < total; cidx += 64) {
    int sb = 0;
    while (cidx >= S.spre[sb + 1]) ++sb;
    const int nl = S.snl[sb], lay = S.slay[sb];
    int idx[4];
    uint32_t rest = (uint32_t)(cidx - S.spre[sb]);
    for (int i = nl - 1; i >= 0; --i) {
      const int l = (lay >> (4 * i)) & 15;
      const uint32_t sz = (uint32_t)(lay_off[l + 1] - lay_off[l]);
      idx[i] = lay_off[l] + (int)(rest % sz);
      rest /= sz;
    }
    bool skip = true;
    double mdt = -INFINITY, lo = INFINITY, hi = -INFINITY, msd = -INFINITY;
    double tmin = INFINITY, tmax = -INFINITY;
    for (int i = 0; i < nl; ++i) {
      const int h = S.byl[idx[i]];
      const Cut c = UNI ? cu : cut_of(t, S.sec[h]);
      if (!((int)c.coincidence_layers > nl)) skip = false;
      mdt = fmax(mdt, c.max_dt);
      lo = fmin(lo, c.min_slope);
      hi = fmax(hi, c.max_slope);
      msd = fmax(msd, c.max_slope_diff);
      tmin = fmin(tmin, S.t[h]);
      tmax = fmax(tmax, S.t[h]);
    }
    bool pass = !skip && !(tmax - tmin > mdt);
    double sl[3];
    for (int i = 0; pass && i + 1 < nl; ++i) {
      const int a = S.byl[idx[i]], b = S.byl[idx[i + 1]];
      const double v = (S.x[b] - S.x[a]) / (S.y[b] - S.y[a]);
      if (v < lo || v > hi) pass = false;
      sl[i] = v;
    }
    if (pass && nl > 2 && fabs(sl[0] - sl[1]) > msd) pass = false;
    if (pass && nl > 3 && (fabs(sl[0] - sl[2]) > msd || fabs(sl[1] - sl[2]) > msd)) pass = false;
    if (pass) {
      int ord[4];
      for (int i = 0; i < nl; ++i) ord[i] = i;
      for (int i = 1; i < nl; ++i) {  // stable insertion sort by x
        const int v = ord[i];
        const double xv = S.x[S.byl[idx[v]]];
        int j = i - 1;
        while (j >= 0 && S.x[S.byl[idx[ord[j]]]] > xv) {
          ord[j + 1] = ord[j];
          --j;
        }
        ord[j + 1] = v;
      }
      for (int i = 0; i < nl; ++i) {
        const int q = idx[ord[i]];
        const double x = S.x[S.byl[q]];
        double w = -INFINITY;
        if (i > 0) w = fmax(w, x - S.x[S.byl[idx[ord[i - 1]]]]);
        if (i + 1 < nl) w = fmax(w, S.x[S.byl[idx[ord[i + 1]]]] - x);
        atomicMax(&S.wq[q], w);  // LDS floating max (the distances are finite)
        S.kept[q] = 1;
      }
    }
  }
}

// checkCoincidence of one cluster (mem[0..n), reference reco/coinc.py:184-207): its kept hits appended to the
// event's coincidence list in pulse order with their widened distance; returns the new list length. A cluster whose
// hits share one sector (the usual case) reads that sector's cuts once.
template <int WMAX>
__device__ int wave_check(const crv_params &p, const Tables &t, WS<WMAX> &S, const uint8_t *mem, int n, int nco) {
  constexpr int WNW = WMAX / 64;
  const int lane = threadIdx.x;
  if (n == 0) return nco;
  const int sec0 = S.sec[mem[0]];
  bool same = true;
  for (int i = lane; i < n; i += 64) same = same && S.sec[mem[i]] == sec0;
  const bool uni = __ballot(!same) == 0ull;
  Cut cu{};
  int nmin = 1 << 30, nmax = -1;
  if (uni) {
    cu = cut_of(t, sec0);
    nmin = nmax = (int)cu.coincidence_layers;
  } else {
    for (int i = lane; i < n; i += 64) {
      const int cl = (int)cut_of(t, S.sec[mem[i]]).coincidence_layers;
      nmin = min(nmin, cl);
      nmax = max(nmax, cl);
    }
    nmin = wave_min(nmin);
    nmax = wave_max(nmax);
  }
  int lay_off[5];
  int pos = 0;
  for (int l = 0; l < 4; ++l) {
    lay_off[l] = pos;
    for (int base = 0; base < n; base += 64) {
      const int i = base + lane;
      const bool pr = i < n && S.layer[mem[i]] == l;
      const uint64_t bm = __ballot(pr);
      if (pr) {
        const int q = pos + __popcll(bm & lanes_below());
        S.byl[q] = mem[i];
        S.wq[q] = uni ? cu.final_max_distance : cut_of(t, S.sec[mem[i]]).final_max_distance;
        S.kept[q] = 0;
      }
      pos += __popcll(bm);
    }
  }
  lay_off[4] = pos;
  if (lane < WNW) S.kmask[lane] = 0ull;
  __syncthreads();
  int nonempty = 0;
  for (int l = 0; l < 4; ++l) nonempty += lay_off[l + 1] > lay_off[l];
  if (n > p.big_cluster_threshold && nonempty >= nmin) {
    for (int i = lane; i < n; i += 64) {
      S.coinc[nco + i] = mem[i];
      S.cmaxd[nco + i] = S.maxd[mem[i]];
    }
    __syncthreads();
    return nco + n;
  }
  if (uni) wave_combos<WMAX, true>(t, S, lay_off, nmin, nmax, cu);
  else wave_combos<WMAX, false>(t, S, lay_off, nmin, nmax, cu);
  __syncthreads();
  for (int q = lane; q < pos; q += 64)
    if (S.kept[q]) {
      const int h = S.byl[q];
      S.wm[h] = S.wq[q];  // a pulse is in one cluster: the slot by pulse is free
      atomicOr((unsigned long long *)&S.kmask[h >> 6], 1ull << (h & 63));
    }
  __syncthreads();
  // kept hits in pulse order: a hit's rank is the number of kept hits below it
  int before = 0;
  for (int w = 0; w < WNW; ++w) {
    const uint64_t m = S.kmask[w];
    if ((m >> lane) & 1ull) {
      const int r = nco + before + __popcll(m & lanes_below());
      S.coinc[r] = (uint8_t)(w * 64 + lane);
      S.cmaxd[r] = S.wm[w * 64 + lane];
    }
    before += __popcll(m);
  }
  __syncthreads();
  return nco + before;
}

template <int WMAX>
__global__ __launch_bounds__(64) void k_coinc_wave(crv_params p, Tables t, const int32_t *__restrict__ list,
                                                   const int *__restrict__ lcount, const int64_t *__restrict__ eoff,
                                                   const int64_t *__restrict__ evk, const double *__restrict__ ptime,
                                                   const double *__restrict__ ppe, Ev e, int32_t *__restrict__ ncl,
                                                   int32_t *__restrict__ nmem, ClSlots slots,
                                                   int32_t *__restrict__ cl_moff, int32_t *__restrict__ mem,
                                                   int *__restrict__ flags) {
  constexpr int WNW = WMAX / 64;
  __shared__ WS<WMAX> S;
  const int lane = threadIdx.x;
  const int64_t nlist = *lcount;  // the events of this size class (k_block_keys' list)
  for (int64_t li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int64_t g = list[li];
    const int64_t a = eoff[g], n64 = eoff[g + 1] - a;
    if (n64 > WMAX) continue;  // not of this class (never listed so)
    const int n = (int)n64;
    const int64_t evkey = evk[g];
    int32_t *cl_moff_e = cl_moff + a + g;
    int32_t *mem_e = mem + a;
    int nclus = 0, nm = 0;
    for (int k = lane; k < n; k += 64) {
      S.bar[k] = e.bar[a + k];
      S.sipm[k] = (uint8_t)(e.sipm[a + k] & 0xFF);
      S.seq0[k] = e.seq[a + k] == 0;
    }
    __syncthreads();
    for (int ti = 0; ti < t.ntypes; ++ti) {
      const int type = t.type_id[ti];
      const double P0 = t.tp0[ti], P1 = t.tp1[ti], half = t.thalf[ti], tmaxdt = t.tmaxdt[ti];
      const bool both = t.tside0[ti] && t.tside1[ti];
      // hits of this sector type, in pulse order
      int nh = 0;
      for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        bool hit = false;
        if (k < n && !(p.ignore_secondary_pulses && !S.seq0[k])) {
          const int bar = e.bar[a + k], sp = e.sipm[a + k];
          if (bar < 0 || bar >= t.nbars || sp < 0 || sp > 3 || bar * 4 + 3 >= t.n_channels) {
            atomicAdd(&flags[F_BADCHAN], 1);
          } else {
            const BarInfo bi = t.binfo[bar];
            if (bi.type == type) {
              double tt = ptime[a + k];
              if (sp % 2 == 0) tt += (bi.ps0 - P0) / p.fiber_signal_speed;
              else tt += (P1 - bi.ps1) / p.fiber_signal_speed;
              double pe = ppe[a + k];
              const int other = bar * 4 + sp + (sp < 2 ? 2 : -2);
              if (t.status[other] & p.compensate_mask) pe *= 2.0;
              S.t[k] = tt;
              S.pe[k] = pe;
              S.maxd[k] = bi.imd;
              S.x[k] = bi.x;
              S.y[k] = bi.y;
              S.sec[k] = (int16_t)bi.sec;
              S.layer[k] = (uint8_t)bi.layer;
              S.cit[k] = (int16_t)bi.cit;
              S.side[k] = (uint8_t)(sp % 2);
              hit = true;
            }
          }
        }
        const uint64_t bm = __ballot(hit);
        if (hit) S.hl[nh + __popcll(bm & lanes_below())] = (uint8_t)k;
        nh += __popcll(bm);
      }
      __syncthreads();
      if (nh == 0) continue;
      // filterHits (reco/coinc.py:154-182), both sides at once: a hit sums the PE of the hits of its side and layer
      // within max_dt_adjacent, in hit order, by counter offset; the hits of a (side, layer) are listed together
      {
        int pos = 0;
        for (int grp = 0; grp < 8; ++grp) {
          if (lane == 0) S.gof[grp] = (int16_t)pos;
          for (int base = 0; base < nh; base += 64) {
            const int ii = base + lane;
            const bool pr = ii < nh && S.side[S.hl[ii]] * 4 + S.layer[S.hl[ii]] == grp;
            const uint64_t bm = __ballot(pr);
            if (pr) S.sl[pos + __popcll(bm & lanes_below())] = S.hl[ii];
            pos += __popcll(bm);
          }
        }
        if (lane == 0) S.gof[8] = (int16_t)pos;
        __syncthreads();
        for (int ii = lane; ii < nh; ii += 64) {
          const int h = S.hl[ii];
          const Cut ch = cut_of(t, S.sec[h]);
          const int grp = S.side[h] * 4 + S.layer[h], cnh = S.cit[h];
          const double th = S.t[h];
          double same = 0.0, adj1 = 0.0, adj2 = 0.0;
          for (int jj = S.gof[grp]; jj < S.gof[grp + 1]; ++jj) {
            const int q = S.sl[jj];
            if (fabs(S.t[q] - th) > ch.max_dt_adjacent) continue;
            const int d = S.cit[q] - cnh;
            same += d == 0 ? S.pe[q] : 0.0;
            adj1 += d == -1 ? S.pe[q] : 0.0;
            adj2 += d == 1 ? S.pe[q] : 0.0;
          }
          S.fpass[ii] = same + adj1 >= ch.pe_threshold || same + adj2 >= ch.pe_threshold;
        }
        __syncthreads();
      }
      // per side: the passing hits in hit order, their clusters and the coincidence check of each
      int nco = 0, nfilt = 0;
      for (int side = 0; side < 2; ++side) {
        const int f0 = nfilt;
        for (int base = 0; base < nh; base += 64) {
          const int ii = base + lane;
          const bool keep = ii < nh && S.side[S.hl[ii]] == side && S.fpass[ii];
          const uint64_t bm = __ballot(keep);
          if (keep) S.filt[nfilt + __popcll(bm & lanes_below())] = S.hl[ii];
          nfilt += __popcll(bm);
        }
        __syncthreads();
        const int ncs = wave_clusters(t, S, S.filt + f0, nfilt - f0, S.maxd, tmaxdt, false, 0.0, S.items, S.off);
        for (int c = 0; c < ncs; ++c) nco = wave_check(p, t, S, S.items + S.off[c], S.off[c + 1] - S.off[c], nco);
      }
      const double dt_opp = 2.0 * half / p.fiber_signal_speed;
      for (int i = lane; i < nco; i += 64) S.wm[S.coinc[i]] = S.cmaxd[i];
      __syncthreads();
      const int ncl2 = wave_clusters(t, S, S.coinc, nco, S.wm, tmaxdt, both, dt_opp, S.cl2, S.off2);
      const int nitems = S.off2[ncl2];
      // attachment of opposite-side pulses (reco/coinc_cpu.py:11-51), one candidate per lane
      int nchosen = 0;
      if (both && p.add_opposite_side_pulses && nitems > 0 && nfilt > 0) {
        if (lane < WNW) S.mmask[lane] = 0ull;
        __syncthreads();
        for (int m = lane; m < nitems; m += 64)
          atomicOr((unsigned long long *)&S.mmask[S.cl2[m] >> 6], 1ull << (S.cl2[m] & 63));
        __syncthreads();
        const double mdt = tmaxdt + dt_opp;
        for (int base = 0; base < nfilt; base += 64) {
          const int ci = base + lane;
          int best_c = -1;
          double best_dt = 0.0;
          if (ci < nfilt) {
            const int h = S.filt[ci];
            const bool member = (S.mmask[h >> 6] >> (h & 63)) & 1ull;
            if (S.seq0[h] && !member) {
              const int hb = S.bar[h], hs = S.side[h];
              for (int c = 0; c < ncl2; ++c)
                for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) {
                  const int an = S.cl2[m];
                  if (!S.seq0[an] || S.bar[an] != hb || hs != 1 - S.side[an]) continue;
                  const double dt = fabs(S.t[h] - S.t[an]);
                  if (!(dt < mdt)) continue;
                  bool occ = false;
                  for (int q = S.off2[c]; q < S.off2[c + 1] && !occ; ++q)
                    occ = S.bar[S.cl2[q]] == hb && S.side[S.cl2[q]] == hs;
                  if (occ) continue;
                  if (best_c < 0 || dt < best_dt || (dt == best_dt && c < best_c)) {
                    best_c = c;
                    best_dt = dt;
                  }
                }
            }
            S.chk[ci] = (int16_t)best_c;
            S.cdt[ci] = best_dt;
          }
        }
        __syncthreads();
        for (int base = 0; base < nfilt; base += 64) {  // the chosen candidates in candidate order (items is free)
          const int ci = base + lane;
          const bool chosen = ci < nfilt && S.chk[ci] >= 0;
          const uint64_t bm = __ballot(chosen);
          if (chosen) S.items[nchosen + __popcll(bm & lanes_below())] = (uint8_t)ci;
          nchosen += __popcll(bm);
        }
        __syncthreads();
        // order (cluster, bar, sipm, dt, candidate): rank of each chosen candidate among the chosen
        for (int u = lane; u < nchosen; u += 64) {
          const int ci = S.items[u];
          const int k = S.chk[ci], h = S.filt[ci];
          const int b = S.bar[h], sp = S.sipm[h];
          const double dt = S.cdt[ci];
          int rank = 0;
          for (int v = 0; v < nchosen; ++v) {
            const int cj = S.items[v];
            if (cj == ci) continue;
            const int kj = S.chk[cj], hj = S.filt[cj];
            const int bj = S.bar[hj], sj = S.sipm[hj];
            const double dj = S.cdt[cj];
            const bool less = kj < k || (kj == k && (bj < b || (bj == b && (sj < sp || (sj == sp && (dj < dt ||
                              (dj == dt && cj < ci)))))));
            rank += less;
          }
          S.sk[rank] = (int16_t)k;
          S.sc[rank] = (uint8_t)ci;
        }
        __syncthreads();
      }
      // stage the bar position (wq, cmaxd, wm) and min_cluster_pes (cdt) of each hit of this sector type (every
      // cluster member is one; other pulses' sector and bar slots are not valid here): those arrays are done with
      for (int i = lane; i < nh; i += 64) {
        const int h = S.hl[i], b = S.bar[h];
        S.wq[h] = t.bpos[b * 3];
        S.cmaxd[h] = t.bpos[b * 3 + 1];
        S.wm[h] = t.bpos[b * 3 + 2];
        S.cdt[h] = cut_of(t, S.sec[h]).min_cluster_pes;
      }
      __syncthreads();
      // members of each second-level cluster (items, then its attachments, one per (bar, sipm)); min-PE test
      for (int base = 0; base < ncl2; base += 64) {
        const int c = base + lane;
        if (c < ncl2) {
          int mc = S.off2[c + 1] - S.off2[c];
          double tot = 0.0, mincp = INFINITY;
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) {
            const int h = S.cl2[m];
            tot = tot + S.pe[h];
            mincp = fmin(mincp, S.cdt[h]);
          }
          int lastb = -1, lasts = -1;
          for (int q = 0; q < nchosen; ++q) {
            if (S.sk[q] != c) continue;
            const int h = S.filt[S.sc[q]];
            if (S.bar[h] == lastb && S.sipm[h] == lasts) continue;
            lastb = S.bar[h];
            lasts = S.sipm[h];
            ++mc;
            tot = tot + S.pe[h];
            mincp = fmin(mincp, S.cdt[h]);
          }
          S.cl_mc[c] = tot < mincp ? 0 : mc;
        }
      }
      __syncthreads();
      // member offsets of the kept clusters (in cluster order), then members and moments, one cluster per lane
      for (int base = 0; base < ncl2; base += 64) {
        const int c = base + lane;
        const int mc = c < ncl2 ? S.cl_mc[c] : 0;
        int incl = mc, kincl = mc > 0;
        for (int o = 1; o < 64; o <<= 1) {
          const int v = __shfl_up(incl, o), kv = __shfl_up(kincl, o);
          if (lane >= o) {
            incl += v;
            kincl += kv;
          }
        }
        const int mbase = nm + incl - mc, cslot = nclus + kincl - (mc > 0);
        if (mc > 0) {
          int32_t *mm = mem_e + mbase, *ml = S.mlist + mbase;  // output and LDS copy (a pulse is in one cluster)
          int w = 0;
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m, ++w) mm[w] = ml[w] = S.cl2[m];
          int lastb = -1, lasts = -1;
          for (int q = 0; q < nchosen; ++q) {
            if (S.sk[q] != c) continue;
            const int h = S.filt[S.sc[q]];
            if (S.bar[h] == lastb && S.sipm[h] == lasts) continue;
            lastb = S.bar[h];
            lasts = S.sipm[h];
            mm[w] = ml[w] = h;
            ++w;
          }
          cluster_emit(p, t, WaveAcc<WMAX>{S}, ml, mc, type, P0, P1, evkey, a + cslot, slots);
          cl_moff_e[cslot] = mbase;
        }
        nm += __shfl(incl, 63);
        nclus += __shfl(kincl, 63);
      }
      __syncthreads();
    }
    if (lane == 0) {
      cl_moff_e[nclus] = nm;
      ncl[g] = nclus;
      nmem[g] = nm;
    }
    __syncthreads();
  }
}

// ---- block finder: one 256-thread block per event of more than 256 pulses (showers, overlays) ----
// The steps of k_coinc with algorithms that scale with the event: filterHits over (side, layer, counter column)
// buckets, each summed in hit order (the FP64 sums of the serial scan); greedy clusters as the connected components
// of the link graph in the order of their lowest list position, with the join order given by the pass number
//   pass(seed) = 1,  pass(p) = min over linked q of pass(q) + [q > p]
// (a hit joins in the pass after its earliest linked member unless that member comes before it in the list), ties
// by list position; link candidates from time bins at least one link window wide; layer combinations by one thread
// per cluster or by the whole block for large products; attachments looked up by bar.
constexpr int BT = 256;      // threads of a block-finder block
constexpr int BH = 12288;    // LDS histogram bins: (side, layer, counter column) buckets, time bins, bars
constexpr int BIG_COMBOS = 512;
constexpr int SMALL_CLUSTER = 24;  // larger clusters are checked by the whole block
static_assert(SMALL_CLUSTER < 64, "a small cluster's layer offsets are packed in 6 bits");
constexpr int INF_PASS = 1 << 29;
constexpr int CM = 1024, CB = 1024;  // clustering staged in LDS: positions, time bins
constexpr int TN = 1536, TG = 384;   // coincidence check staged in LDS: hits, bar groups

struct BS {  // event-local scratch of the block finder (index: pulse in the event, or position in a list)
  double *t, *pe, *x, *y, *maxd, *wm, *cmaxd, *wq, *cdt, *px, *pt, *pmd;
  int32_t *bar, *sec, *cit, *hl, *filt, *items, *off, *coinc, *byl, *cl2, *off2, *bsort, *tsort, *label, *pass, *chk,
      *sk, *sc, *clmc, *tmp, *clid, *tbin;
  uint8_t *layer, *side, *sipm, *seq0, *kept, *fpass, *memf, *pside;
};

struct ClLds {  // findClusters of up to CM positions
  double x[CM], t[CM], md[CM];
  int32_t label[CM], pass[CM], tsort[CM], tbin[CM], bins[CB];
  uint8_t side[CM];
};

struct TpLds {  // checkCoincidence of up to TN hits in up to TG bar groups
  double ts[TN];
  double gx[TG], gy[TG], gtmin[TG], gtmax[TG], gmdt[TG], glo[TG], ghi[TG], gmsd[TG], gw[TG];
  int32_t key[TN], gs[TG + 1];
  int8_t gcl[TG];
  uint8_t gk[TG];
};

struct BlockLds {
  union {
    hipcub::BlockScan<int, BT>::TempStorage scan;
    hipcub::BlockReduce<double, BT>::TempStorage red;
  } cub;
  union {
    int hist[BH];
    ClLds cl;
    TpLds tp;
  } u;
  double dshare[4];
  int ishare[8];
  int flag;
  int qpos;  // the block's current queue position
};

__device__ __forceinline__ int bscan(BlockLds &L, int v, int &total) {
  int r;
  hipcub::BlockScan<int, BT>(L.cub.scan).ExclusiveSum(v, r, total);
  __syncthreads();
  return r;
}

// exclusive scan of h[0..nb) in place (an LDS array); returns the total
__device__ int bscan_arr(BlockLds &L, int *h, int nb) {
  const int per = (nb + BT - 1) / BT, a = threadIdx.x * per;
  int s = 0;
  for (int i = a; i < a + per && i < nb; ++i) s += h[i];
  int tot;
  int run = bscan(L, s, tot);
  for (int i = a; i < a + per && i < nb; ++i) {
    const int v = h[i];
    h[i] = run;
    run += v;
  }
  __syncthreads();
  return tot;
}

__device__ __forceinline__ int bscan_hist(BlockLds &L, int nb) { return bscan_arr(L, L.u.hist, nb); }

__device__ __forceinline__ double block_min(BlockLds &L, double v) {
  v = hipcub::BlockReduce<double, BT>(L.cub.red).Reduce(v, hipcub::Min());
  __syncthreads();
  if (threadIdx.x == 0) L.dshare[3] = v;
  __syncthreads();
  v = L.dshare[3];
  __syncthreads();
  return v;
}

__device__ __forceinline__ double block_max(BlockLds &L, double v) {
  v = hipcub::BlockReduce<double, BT>(L.cub.red).Reduce(v, hipcub::Max());
  __syncthreads();
  if (threadIdx.x == 0) L.dshare[3] = v;
  __syncthreads();
  v = L.dshare[3];
  __syncthreads();
  return v;
}

// the cell grid of findClusters over times [lo, hi] and x [xlo, xhi] (largest link distance mdm): cells at least one
// link window (max_dt, plus dt_opp across sides when `opposite`) wide in time and the largest link distance wide in
// x, so that linked positions lie in neighbouring cells; at most maxb - 2 cells
struct CellGrid {
  double lo, width, xlo, xw;
  bool fin, xfin;
  int nbt, nbx;
  __device__ int cell(double t, double x) const {
    const double ft = fin ? (t - lo) / width : 0.0;
    const int bt = ft >= 0.0 ? min((int)fmin(ft, (double)(nbt - 1)), nbt - 1) : 0;  // NaN -> bin 0
    const double fx = xfin ? (x - xlo) / xw : 0.0;
    const int bx = fx >= 0.0 ? min((int)fmin(fx, (double)(nbx - 1)), nbx - 1) : 0;
    return bt * nbx + bx;
  }
};

__device__ CellGrid cell_grid(double lo, double hi, double xlo, double xhi, double mdm, double max_dt, bool opposite,
                              double dt_opp, int maxb) {
  CellGrid g;
  const double mt = max_dt + (opposite ? dt_opp : 0.0);
  const double xspan = xhi - xlo;
  g.xfin = xspan >= 0.0 && xspan < 1e300 && mdm >= 0.0 && mdm < 1e300;
  const double xw0 = g.xfin ? fmax(mdm, 1e-9) * (1.0 + 1e-12) + 1e-9 : 1.0;
  g.nbx = g.xfin ? max(1, min((int)(xspan / xw0) + 1, 32)) : 1;
  g.xw = g.xfin ? fmax(xw0, xspan / (double)max(g.nbx - 1, 1) * (1.0 + 1e-12)) : 1.0;
  g.xlo = xlo;
  const int tcap = (maxb - 2) / g.nbx;
  const double span = hi - lo;
  g.fin = span >= 0.0 && span < 1e300;  // else one time bin
  g.width = g.fin ? fmax(mt, span / (double)max(tcap - 1, 1)) * (1.0 + 1e-12) + 1e-9 : 1.0;
  g.nbt = g.fin ? max(1, min((int)(span / g.width) + 1, tcap)) : 1;
  g.lo = lo;
  return g;
}

// the clusters of list[0..m) in seed order, from each position's label (its component's lowest position) and pass:
// member offsets to off[], members (pulses) to out[] by (pass, position); ITM: m ints of scratch
__device__ int components_out(BlockLds &L, const BS &S, const int32_t *list, int m, const int32_t *LAB,
                              const int32_t *PAS, int32_t *ITM, int32_t *out, int32_t *off) {
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += BT) S.tmp[i] = 0;
  __syncthreads();
  for (int i = tid; i < m; i += BT) atomicAdd(&S.tmp[LAB[i]], 1);
  __syncthreads();
  int base = 0, cbase = 0;
  for (int c = 0; c < m; c += BT) {  // member offsets and cluster numbers over seeds in position order
    const int i = c + tid;
    const int cnt = i < m ? S.tmp[i] : 0;
    int tot, ctot;
    const int r = bscan(L, cnt, tot);
    const int cr = bscan(L, cnt > 0, ctot);
    if (cnt > 0) {
      off[cbase + cr] = base + r;
      S.clid[i] = cbase + cr;
      S.tmp[i] = base + r;  // cursor
    }
    base += tot;
    cbase += ctot;
  }
  __syncthreads();
  if (tid == 0) off[cbase] = base;
  for (int i = tid; i < m; i += BT) ITM[atomicAdd(&S.tmp[LAB[i]], 1)] = i;  // positions, unordered per cluster
  __syncthreads();
  for (int k = tid; k < base; k += BT) {  // rank by (pass, position) within the cluster
    const int i = ITM[k], c = S.clid[LAB[i]], pi = PAS[i];
    const int o = off[c], e = off[c + 1];
    int r = 0;
    for (int q = o; q < e; ++q) {
      const int j = ITM[q], pj = PAS[j];
      r += pj < pi || (pj == pi && j < i);
    }
    out[o + r] = list[i];
  }
  __syncthreads();
  return cbase;
}

// greedy findClusters of the pulses list[0..m) (reco/coinc.py:125-152): members (pulses) to out[] in join order,
// cluster starts to off[]; returns the number of clusters. Clusters are the link graph's connected components in
// the order of their first position; a member's join order is (pass, position) with pass(seed) = 1 and
// pass(p) = min over linked q of pass(q) + [q > p]. Both are least fixpoints, reached in any order of visits. Link
// candidates come from the neighbouring cells of a (time, x) grid (CellGrid).
//
// Up to CM positions: times, x, link distances and sides staged in LDS in list order.
__device__ int block_clusters_lds(BlockLds &L, const BS &S, const int32_t *list, int m, const double *md,
                                  double max_dt, bool opposite, double dt_opp, int32_t *out, int32_t *off) {
  const int tid = threadIdx.x;
  double *X = L.u.cl.x, *T = L.u.cl.t, *MD = L.u.cl.md;
  uint8_t *SD = L.u.cl.side;
  int32_t *LAB = L.u.cl.label, *PAS = L.u.cl.pass, *TSO = L.u.cl.tsort, *TBI = L.u.cl.tbin, *BINS = L.u.cl.bins;
  double lo = INFINITY, hi = -INFINITY;
  for (int i = tid; i < m; i += BT) {
    const int h = list[i];
    const double tt = S.t[h];
    X[i] = S.x[h];
    T[i] = tt;
    MD[i] = md[h];
    SD[i] = S.side[h];
    lo = fmin(lo, tt);
    hi = fmax(hi, tt);
  }
  lo = block_min(L, lo);
  hi = block_max(L, hi);
  double xlo = INFINITY, xhi = -INFINITY, mdm = -INFINITY;
  for (int i = tid; i < m; i += BT) {
    xlo = fmin(xlo, X[i]);
    xhi = fmax(xhi, X[i]);
    mdm = fmax(mdm, MD[i]);
  }
  xlo = block_min(L, xlo);
  xhi = block_max(L, xhi);
  mdm = block_max(L, mdm);
  const CellGrid G = cell_grid(lo, hi, xlo, xhi, mdm, max_dt, opposite, dt_opp, CB);
  const int nbt = G.nbt, nbx = G.nbx, nb = nbt * nbx;
  for (int b = tid; b < nb + 1; b += BT) BINS[b] = 0;
  __syncthreads();
  for (int i = tid; i < m; i += BT) {
    TBI[i] = G.cell(T[i], X[i]);
    atomicAdd(&BINS[TBI[i]], 1);
  }
  __syncthreads();
  bscan_arr(L, BINS, nb + 1);
  for (int i = tid; i < m; i += BT) TSO[atomicAdd(&BINS[TBI[i]], 1)] = i;  // BINS[b] ends as the end of cell b
  __syncthreads();
  auto link = [&](int i, int j) {
    const double mdd = fmax(MD[i], MD[j]);
    const double w = max_dt + ((opposite && SD[i] != SD[j]) ? dt_opp : 0.0);
    return fabs(X[i] - X[j]) <= mdd && fabs(T[i] - T[j]) < w;
  };
  // components: min-label propagation with pointer jumping
  for (int i = tid; i < m; i += BT) {
    LAB[i] = i;
    PAS[i] = INF_PASS;
  }
  __syncthreads();
  for (;;) {  // every thread leaves together: the vote is a barrier with a block-wide OR
    int changed = 0;
    for (int i = tid; i < m; i += BT) {
      const int bt = TBI[i] / nbx, bx = TBI[i] - bt * nbx;
      int best = LAB[i];
      for (int r = max(bt - 1, 0); r <= min(bt + 1, nbt - 1); ++r) {
        const int c0 = r * nbx + max(bx - 1, 0), c1 = r * nbx + min(bx + 1, nbx - 1);
        for (int k = c0 ? BINS[c0 - 1] : 0; k < BINS[c1]; ++k) {
          const int j = TSO[k];
          const int lj = LAB[j];
          if (lj < best && j != i && link(i, j)) best = lj;
        }
      }
      if (best < LAB[i]) {
        atomicMin(&LAB[i], best);
        changed = 1;
      }
    }
    __syncthreads();
    for (int i = tid; i < m; i += BT) {
      int l = LAB[i];
      while (LAB[l] < l) l = LAB[l];
      LAB[i] = l;
    }
    __syncthreads();
    if (!__syncthreads_or(changed)) break;
  }
  // pass numbers: 0-1 relaxation from the seeds (a component's lowest position)
  for (int i = tid; i < m; i += BT)
    if (LAB[i] == i) PAS[i] = 1;
  __syncthreads();
  for (;;) {
    int changed = 0;
    for (int i = tid; i < m; i += BT) {
      if (PAS[i] == 1) continue;  // a seed or a first-pass member: no lower pass exists
      const int bt = TBI[i] / nbx, bx = TBI[i] - bt * nbx;
      int best = PAS[i];
      for (int r = max(bt - 1, 0); r <= min(bt + 1, nbt - 1); ++r) {
        const int c0 = r * nbx + max(bx - 1, 0), c1 = r * nbx + min(bx + 1, nbx - 1);
        for (int k = c0 ? BINS[c0 - 1] : 0; k < BINS[c1]; ++k) {
          const int j = TSO[k];
          const int cand = PAS[j] + (j > i ? 1 : 0);
          if (cand < best && j != i && link(i, j)) best = cand;
        }
      }
      if (best < PAS[i]) {
        PAS[i] = best;
        changed = 1;
      }
    }
    __syncthreads();
    if (!__syncthreads_or(changed)) break;
  }
  return components_out(L, S, list, m, LAB, PAS, TBI, out, off);  // the bins are no longer needed
}

// union-find over cell-order positions (par in LDS): a root hooks under the root of the lower list position, so the
// root of every component ends as its lowest list position
__device__ __forceinline__ int uf_find(const volatile int *par, int x) {
  int y;
  while ((y = par[x]) != x) x = y;
  return x;
}
__device__ void uf_unite(int *par, const int32_t *tso, int a, int b) {
  const volatile int *vp = par;
  for (;;) {
    a = uf_find(vp, a);
    b = uf_find(vp, b);
    if (a == b) return;
    if (tso[a] < tso[b]) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(par + a, a, b) == a) return;  // else a was hooked meanwhile: again from its new root
  }
}

// Longer lists: the positions' data gathered in cell order into global scratch (px / pt / pmd / pside; the
// cell-order labels in tbin, passes in clid, cells in items; the inverse order in tmp), so that the candidates of
// neighbouring positions are neighbouring memory that the wavefront reads once; the fixpoints come out in list order
// in label / pass. Each position's links are found once: a bit per candidate of its three rows of cells, in row
// order, kept as a 64-bit mask in cdt (free while clusters are formed); the rounds then visit linked candidates
// only. A position with more than 63 candidates has bit 63 set and tests its links in every round. While the bins
// leave room for two ints per position in the LDS histogram (up to ~5,600 positions), the labels come from a
// union-find over the links (one pass) and the pass rounds run on LDS; longer lists propagate labels in rounds.
__device__ int block_clusters_cells(BlockLds &L, const BS &S, const int32_t *list, int m, const double *md,
                                    double max_dt, bool opposite, double dt_opp, int32_t *out, int32_t *off) {
  const int tid = threadIdx.x;
  double *X = S.px, *T = S.pt, *MD = S.pmd;  // cell order
  uint8_t *SD = S.pside;
  int32_t *TSO = S.tsort;  // cell order -> list position
  int32_t *TBI = S.tbin;   // list position -> cell, then the cell-order labels
  int32_t *CEL = S.items;  // cell order -> cell
  int32_t *RNK = S.tmp;    // list position -> cell order
  int32_t *LABC = S.tbin, *PASC = S.clid;
  int32_t *BINS = L.u.hist;
  double lo = INFINITY, hi = -INFINITY, xlo = INFINITY, xhi = -INFINITY, mdm = -INFINITY;
  for (int i = tid; i < m; i += BT) {
    const int h = list[i];
    const double tt = S.t[h], xx = S.x[h];
    lo = fmin(lo, tt);
    hi = fmax(hi, tt);
    xlo = fmin(xlo, xx);
    xhi = fmax(xhi, xx);
    mdm = fmax(mdm, md[h]);
  }
  lo = block_min(L, lo);
  hi = block_max(L, hi);
  xlo = block_min(L, xlo);
  xhi = block_max(L, xhi);
  mdm = block_max(L, mdm);
  const bool lds_rounds = 2 * m <= BH - CB;  // parents and passes beside at most BH - 2 m bins
  const CellGrid G = cell_grid(lo, hi, xlo, xhi, mdm, max_dt, opposite, dt_opp, lds_rounds ? BH - 2 * m : BH);
  const int nbt = G.nbt, nbx = G.nbx, nb = nbt * nbx;
  for (int b = tid; b < nb + 1; b += BT) BINS[b] = 0;
  __syncthreads();
  for (int i = tid; i < m; i += BT) {
    const int h = list[i];
    TBI[i] = G.cell(S.t[h], S.x[h]);
    atomicAdd(&BINS[TBI[i]], 1);
  }
  __syncthreads();
  bscan_arr(L, BINS, nb + 1);
  for (int i = tid; i < m; i += BT) TSO[atomicAdd(&BINS[TBI[i]], 1)] = i;  // BINS[b] ends as the end of cell b
  __syncthreads();
  for (int k = tid; k < m; k += BT) {
    const int i = TSO[k], h = list[i];
    X[k] = S.x[h];
    T[k] = S.t[h];
    MD[k] = md[h];
    SD[k] = S.side[h];
    CEL[k] = TBI[i];
    RNK[i] = k;
  }
  __syncthreads();
  for (int k = tid; k < m; k += BT) {
    LABC[k] = TSO[k];
    PASC[k] = INF_PASS;
  }
  __syncthreads();
  auto link = [&](int a, int b) {
    const double mdd = fmax(MD[a], MD[b]);
    const double w = max_dt + ((opposite && SD[a] != SD[b]) ? dt_opp : 0.0);
    return fabs(X[a] - X[b]) <= mdd && fabs(T[a] - T[b]) < w;
  };
  constexpr uint64_t OVER = 1ull << 63;
  uint64_t *ADJ = reinterpret_cast<uint64_t *>(S.cdt);
  // fn(k2) for every candidate k2 of position k linked to it (cell-order indices)
  auto linked = [&](int k, uint64_t adj, auto &&fn) {
    const int bt = CEL[k] / nbx, bx = CEL[k] - bt * nbx;
    int q = 0;
    for (int r = max(bt - 1, 0); r <= min(bt + 1, nbt - 1); ++r) {
      const int c0 = r * nbx + max(bx - 1, 0), c1 = r * nbx + min(bx + 1, nbx - 1);
      const int s0 = c0 ? BINS[c0 - 1] : 0, s1 = BINS[c1];
      if (adj & OVER) {
        for (int k2 = s0; k2 < s1; ++k2)
          if (k2 != k && link(k, k2)) fn(k2);
      } else {
        for (uint64_t b = q < 63 ? (adj >> q) & ((s1 - s0 >= 64 ? ~0ull : (1ull << (s1 - s0)) - 1ull)) : 0ull; b;
             b &= b - 1)
          fn(s0 + __ffsll((long long)b) - 1);
      }
      q += s1 - s0;
    }
  };
  for (int k = tid; k < m; k += BT) {
    const int bt = CEL[k] / nbx, bx = CEL[k] - bt * nbx;
    uint64_t adj = 0;
    int q = 0;
    for (int r = max(bt - 1, 0); r <= min(bt + 1, nbt - 1); ++r) {
      const int c0 = r * nbx + max(bx - 1, 0), c1 = r * nbx + min(bx + 1, nbx - 1);
      for (int k2 = c0 ? BINS[c0 - 1] : 0; k2 < BINS[c1] && !(adj & OVER); ++k2, ++q) {
        if (q >= 63) adj = OVER;
        else if (k2 != k && link(k, k2)) adj |= 1ull << q;
      }
    }
    ADJ[k] = adj;
  }
  __syncthreads();
  if (lds_rounds) {
    int *PAR = BINS + nb + 1, *PASL = PAR + m;
    for (int k = tid; k < m; k += BT) PAR[k] = k;
    __syncthreads();
    for (int k = tid; k < m; k += BT) {
      const uint64_t adj = ADJ[k];
      if (adj) linked(k, adj, [&](int k2) {
          if (k2 > k) uf_unite(PAR, TSO, k, k2);  // each link once
        });
    }
    __syncthreads();
    for (int k = tid; k < m; k += BT) {
      const int r = uf_find(PAR, k);
      LABC[k] = TSO[r];
      PASL[k] = r == k ? 1 : INF_PASS;  // the seed of its component
    }
    __syncthreads();
    for (;;) {  // passes
      int changed = 0;
      for (int k = tid; k < m; k += BT) {
        if (PASL[k] == 1) continue;  // a seed, or already in the first pass: no lower pass exists
        const int i = TSO[k];
        int best = PASL[k];
        linked(k, ADJ[k], [&](int k2) { best = min(best, PASL[k2] + (TSO[k2] > i ? 1 : 0)); });
        if (best < PASL[k]) {
          PASL[k] = best;
          changed = 1;
        }
      }
      __syncthreads();
      if (!__syncthreads_or(changed)) break;
    }
    for (int k = tid; k < m; k += BT) {  // back to list order
      S.label[TSO[k]] = LABC[k];
      S.pass[TSO[k]] = PASL[k];
    }
    __syncthreads();
    return components_out(L, S, list, m, S.label, S.pass, S.items, out, off);
  }
  for (;;) {  // labels: every thread leaves together (the vote is a barrier with a block-wide OR)
    int changed = 0;
    for (int k = tid; k < m; k += BT) {
      const uint64_t adj = ADJ[k];
      if (!adj) continue;
      int best = LABC[k];
      linked(k, adj, [&](int k2) { best = min(best, LABC[k2]); });
      if (best < LABC[k]) {
        atomicMin(&LABC[k], best);
        changed = 1;
      }
    }
    __syncthreads();
    for (int k = tid; k < m; k += BT) {  // pointer jumping over positions' labels
      int l = LABC[k];
      while (LABC[RNK[l]] < l) l = LABC[RNK[l]];
      LABC[k] = l;
    }
    __syncthreads();
    if (!__syncthreads_or(changed)) break;
  }
  for (int k = tid; k < m; k += BT)
    if (LABC[k] == TSO[k]) PASC[k] = 1;
  __syncthreads();
  for (;;) {  // passes
    int changed = 0;
    for (int k = tid; k < m; k += BT) {
      const int i = TSO[k];
      if (LABC[k] == i) continue;
      int best = PASC[k];
      linked(k, ADJ[k], [&](int k2) { best = min(best, PASC[k2] + (TSO[k2] > i ? 1 : 0)); });
      if (best < PASC[k]) {
        PASC[k] = best;
        changed = 1;
      }
    }
    __syncthreads();
    if (!__syncthreads_or(changed)) break;
  }
  for (int k = tid; k < m; k += BT) {  // back to list order
    S.label[TSO[k]] = LABC[k];
    S.pass[TSO[k]] = PASC[k];
  }
  __syncthreads();
  return components_out(L, S, list, m, S.label, S.pass, S.items, out, off);
}

__device__ int block_clusters(BlockLds &L, const BS &S, const int32_t *list, int m, const double *md, double max_dt,
                              bool opposite, double dt_opp, int32_t *out, int32_t *off) {
  if (m == 0) {
    if (threadIdx.x == 0) off[0] = 0;
    __syncthreads();
    return 0;
  }
  return m <= CM ? block_clusters_lds(L, S, list, m, md, max_dt, opposite, dt_opp, out, off)
                 : block_clusters_cells(L, S, list, m, md, max_dt, opposite, dt_opp, out, off);
}

// checkCoincidence of a cluster's members mem[0..n) for one thread; by-layer lists in byl[0..n) with their widened
// distance in wq / kept flags (same offsets); returns the number of combinations when they are too many for one
// thread (the block then enumerates them), else 0
__device__ int64_t thread_check(const crv_params &p, const Tables &t, const BS &S, const int32_t *mem, int n,
                                int32_t *byl, double *wq, uint8_t *kept, int lay_off[5], int &nmin, int &nmax,
                                bool &all_kept) {
  nmin = 1 << 30;
  nmax = -1;
  for (int i = 0; i < n; ++i) {
    const int cl = (int)cut_of(t, S.sec[mem[i]]).coincidence_layers;
    nmin = min(nmin, cl);
    nmax = max(nmax, cl);
  }
  int pos = 0;
  for (int l = 0; l < 4; ++l) {
    lay_off[l] = pos;
    for (int i = 0; i < n; ++i)
      if (S.layer[mem[i]] == l) {
        byl[pos] = mem[i];
        wq[pos] = cut_of(t, S.sec[mem[i]]).final_max_distance;
        kept[pos] = 0;
        ++pos;
      }
  }
  lay_off[4] = pos;
  int nonempty = 0;
  for (int l = 0; l < 4; ++l) nonempty += lay_off[l + 1] > lay_off[l];
  all_kept = n > p.big_cluster_threshold && nonempty >= nmin;
  if (all_kept) return 0;
  int64_t total = 0;
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int Lc[4];
    for (int i = 0; i < nl; ++i) Lc[i] = i;
    for (;;) {
      int64_t P = 1;
      for (int i = 0; i < nl; ++i) P *= lay_off[Lc[i] + 1] - lay_off[Lc[i]];
      total += P;
      int k = nl - 1;
      while (k >= 0 && Lc[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++Lc[k];
      for (int i = k + 1; i < nl; ++i) Lc[i] = Lc[i - 1] + 1;
    }
  }
  return total;
}

// layer offsets of a cluster's by-layer lists and the range of its coincidence_layers (byl already built)
__device__ void cluster_layout(const Tables &t, const BS &S, const int32_t *mem, int n, int lay_off[5], int &nmin,
                               int &nmax) {
  nmin = 1 << 30;
  nmax = -1;
  int pos = 0;
  for (int i = 0; i < n; ++i) {
    const int cl = (int)cut_of(t, S.sec[mem[i]]).coincidence_layers;
    nmin = min(nmin, cl);
    nmax = max(nmax, cl);
  }
  for (int l = 0; l < 4; ++l) {
    lay_off[l] = pos;
    for (int i = 0; i < n; ++i) pos += S.layer[mem[i]] == l;
  }
  lay_off[4] = pos;
}

// one layer combination idx[0..nl) of byl: the pass test, then widening and kept flags (atomic when shared)
__device__ __forceinline__ void combo_eval(const Tables &t, const BS &S, const int32_t *byl, double *wq, uint8_t *kept,
                                           const int idx[4], int nl, bool atomic_) {
  bool skip = true;
  double mdt = -INFINITY, lo = INFINITY, hi = -INFINITY, msd = -INFINITY, tmin = INFINITY, tmax = -INFINITY;
  for (int i = 0; i < nl; ++i) {
    const int h = byl[idx[i]];
    const Cut c = cut_of(t, S.sec[h]);
    if (!((int)c.coincidence_layers > nl)) skip = false;
    mdt = fmax(mdt, c.max_dt);
    lo = fmin(lo, c.min_slope);
    hi = fmax(hi, c.max_slope);
    msd = fmax(msd, c.max_slope_diff);
    tmin = fmin(tmin, S.t[h]);
    tmax = fmax(tmax, S.t[h]);
  }
  bool pass = !skip && !(tmax - tmin > mdt);
  double sl[3];
  for (int i = 0; pass && i + 1 < nl; ++i) {
    const int a = byl[idx[i]], b = byl[idx[i + 1]];
    const double v = (S.x[b] - S.x[a]) / (S.y[b] - S.y[a]);
    if (v < lo || v > hi) pass = false;
    sl[i] = v;
  }
  if (pass && nl > 2 && fabs(sl[0] - sl[1]) > msd) pass = false;
  if (pass && nl > 3 && (fabs(sl[0] - sl[2]) > msd || fabs(sl[1] - sl[2]) > msd)) pass = false;
  if (!pass) return;
  int ord[4];
  for (int i = 0; i < nl; ++i) ord[i] = i;
  for (int i = 1; i < nl; ++i) {
    const int v = ord[i];
    const double xv = S.x[byl[idx[v]]];
    int j = i - 1;
    while (j >= 0 && S.x[byl[idx[ord[j]]]] > xv) {
      ord[j + 1] = ord[j];
      --j;
    }
    ord[j + 1] = v;
  }
  for (int i = 0; i < nl; ++i) {
    const int q = idx[ord[i]];
    const double x = S.x[byl[q]];
    double w = -INFINITY;
    if (i > 0) w = fmax(w, x - S.x[byl[idx[ord[i - 1]]]]);
    if (i + 1 < nl) w = fmax(w, S.x[byl[idx[ord[i + 1]]]] - x);
    if (atomic_) {
      unsigned long long *a = (unsigned long long *)&wq[q];
      unsigned long long old = *a, assumed;
      do {
        assumed = old;
        if (!(w > __longlong_as_double((long long)assumed))) break;
        old = atomicCAS(a, assumed, (unsigned long long)__double_as_longlong(w));
      } while (old != assumed);
    } else {
      wq[q] = fmax(wq[q], w);
    }
    kept[q] = 1;
  }
}

// every combination of the cluster's by-layer lists from combination number c0 with stride cs
__device__ void combos(const Tables &t, const BS &S, const int32_t *byl, double *wq, uint8_t *kept, const int lay_off[5],
                       int nmin, int nmax, int64_t c0, int64_t cs, bool atomic_) {
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int Lc[4];
    for (int i = 0; i < nl; ++i) Lc[i] = i;
    for (;;) {
      int64_t P = 1;
      for (int i = 0; i < nl; ++i) P *= lay_off[Lc[i] + 1] - lay_off[Lc[i]];
      for (int64_t ci = c0; ci < P; ci += cs) {
        int idx[4];
        int64_t rest = ci;
        for (int i = nl - 1; i >= 0; --i) {
          const int sz = lay_off[Lc[i] + 1] - lay_off[Lc[i]];
          idx[i] = lay_off[Lc[i]] + (int)(rest % sz);
          rest /= sz;
        }
        combo_eval(t, S, byl, wq, kept, idx, nl, atomic_);
      }
      int k = nl - 1;
      while (k >= 0 && Lc[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++Lc[k];
      for (int i = k + 1; i < nl; ++i) Lc[i] = Lc[i - 1] + 1;
    }
  }
}

// combination li of a cluster's by-layer lists in combos()'s order (layer subsets by size, then lexicographic):
// its by-layer positions idx[0..nl); returns nl (0 past the last combination)
__device__ __forceinline__ int combo_at(int li, const int lay_off[5], int nmin, int nmax, int idx[4]) {
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int Lc[4];
    for (int i = 0; i < nl; ++i) Lc[i] = i;
    for (;;) {
      int P = 1;
      for (int i = 0; i < nl; ++i) P *= lay_off[Lc[i] + 1] - lay_off[Lc[i]];
      if (li < P) {
        for (int i = nl - 1; i >= 0; --i) {
          const int sz = lay_off[Lc[i] + 1] - lay_off[Lc[i]];
          idx[i] = lay_off[Lc[i]] + li % sz;
          li /= sz;
        }
        return nl;
      }
      li -= P;
      int k = nl - 1;
      while (k >= 0 && Lc[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++Lc[k];
      for (int i = k + 1; i < nl; ++i) Lc[i] = Lc[i - 1] + 1;
    }
  }
  return 0;
}

// atomic max of a double in global memory (no-op unless w is larger; NaN never stored)
__device__ __forceinline__ void gmax_double(double *a, double w) {
  unsigned long long *q = (unsigned long long *)a;
  unsigned long long old = *q, assumed;
  do {
    assumed = old;
    if (!(w > __longlong_as_double((long long)assumed))) break;
    old = atomicCAS(q, assumed, (unsigned long long)__double_as_longlong(w));
  } while (old != assumed);
}

// A cluster's hits grouped by bar (layer, then bar, then time): every hit of a group has the bar's x, y and sector
// cuts, so the slope tests, the skip test and the widened distances of a layer combination depend only on the groups
// it takes; only its time window depends on the hits. TupL reads groups staged in LDS, TupG reads them through the
// hit arrays in global memory.
struct TupL {
  TpLds &T;
  __device__ double x(int g) const { return T.gx[g]; }
  __device__ double y(int g) const { return T.gy[g]; }
  __device__ double tmin(int g) const { return T.gtmin[g]; }
  __device__ double tmax(int g) const { return T.gtmax[g]; }
  __device__ void cut(int g, int &cl, double &mdt, double &lo, double &hi, double &msd) const {
    cl = T.gcl[g], mdt = T.gmdt[g], lo = T.glo[g], hi = T.ghi[g], msd = T.gmsd[g];
  }
  __device__ int h0(int g) const { return T.gs[g]; }
  __device__ int h1(int g) const { return T.gs[g + 1]; }
  __device__ double ht(int k) const { return T.ts[k]; }
  __device__ void mark(int g, double w) const {
    gmax_double(&T.gw[g], w);
    T.gk[g] = 1;
  }
};

struct TupG {
  const BS &S;
  const Tables &t;
  const int32_t *byl, *gs;
  double *gw;
  uint8_t *gk;
  __device__ int rep(int g) const { return byl[gs[g]]; }
  __device__ double x(int g) const { return S.x[rep(g)]; }
  __device__ double y(int g) const { return S.y[rep(g)]; }
  __device__ double tmin(int g) const { return S.t[rep(g)]; }
  __device__ double tmax(int g) const { return S.t[byl[gs[g + 1] - 1]]; }
  __device__ void cut(int g, int &cl, double &mdt, double &lo, double &hi, double &msd) const {
    const Cut c = cut_of(t, S.sec[rep(g)]);
    cl = (int)c.coincidence_layers, mdt = c.max_dt, lo = c.min_slope, hi = c.max_slope, msd = c.max_slope_diff;
  }
  __device__ int h0(int g) const { return gs[g]; }
  __device__ int h1(int g) const { return gs[g + 1]; }
  __device__ double ht(int k) const { return S.t[byl[k]]; }
  __device__ void mark(int g, double w) const {
    gmax_double(&gw[g], w);
    gk[g] = 1;
  }
};

// the first time >= m in group g is within mdt of m (times ascend within a group)
template <class A>
__device__ __forceinline__ bool tup_covers(const A &G, int g, double m, double mdt) {
  for (int k = G.h0(g); k < G.h1(g); ++k) {
    const double u = G.ht(k);
    if (u >= m) return !(u - m > mdt);
  }
  return false;
}

// whether hit time th of tuple member i takes part in a combination of one hit per group within the time window:
// a combination passes iff every hit is within mdt of its earliest hit, which is th itself or a hit u < th of
// another group (then every other group needs a hit in [u, u + mdt]; the same rounding as max - min)
template <class A>
__device__ bool tup_time_ok(const A &G, const int *g, int nl, int i, double th, double mdt) {
  bool ok = true;
  for (int j = 0; j < nl && ok; ++j)
    if (j != i) ok = tup_covers(G, g[j], th, mdt);
  if (ok) return true;
  for (int j = 0; j < nl; ++j) {
    if (j == i) continue;
    for (int k = G.h0(g[j]); k < G.h1(g[j]); ++k) {
      const double u = G.ht(k);
      if (!(u < th)) break;
      if (th - u > mdt) continue;
      bool all = true;
      for (int j2 = 0; j2 < nl && all; ++j2)
        if (j2 != i && j2 != j) all = tup_covers(G, g[j2], u, mdt);
      if (all) return true;
    }
  }
  return false;
}

// one tuple of groups g[0..nl) in layer order with its consecutive slopes sl[0..nl-1): the combination tests of
// reference reco/coinc.py:156-181 with the tuple's cuts, then the widening of every hit that takes part
template <class A>
__device__ __forceinline__ void tup_leaf(const A &G, const int *g, int nl, const double *sl, double *wq, uint8_t *kept) {
  bool skip = true;
  double mdt = -INFINITY, lo = INFINITY, hi = -INFINITY, msd = -INFINITY, m_all = INFINITY, M_all = -INFINITY;
  double xs[4];
  for (int i = 0; i < nl; ++i) {
    int cl;
    double cmdt, clo, chi, cmsd;
    G.cut(g[i], cl, cmdt, clo, chi, cmsd);
    if (!(cl > nl)) skip = false;
    mdt = fmax(mdt, cmdt);
    lo = fmin(lo, clo);
    hi = fmax(hi, chi);
    msd = fmax(msd, cmsd);
    xs[i] = G.x(g[i]);
    m_all = fmin(m_all, G.tmin(g[i]));
    M_all = fmax(M_all, G.tmax(g[i]));
  }
  if (skip) return;
  for (int i = 0; i + 1 < nl; ++i)
    if (sl[i] < lo || sl[i] > hi) return;
  if (nl > 2 && fabs(sl[0] - sl[1]) > msd) return;
  if (nl > 3 && (fabs(sl[0] - sl[2]) > msd || fabs(sl[1] - sl[2]) > msd)) return;
  int ord[4];
  for (int i = 0; i < nl; ++i) ord[i] = i;
  for (int i = 1; i < nl; ++i) {  // stable by x: ties keep layer order
    const int v = ord[i];
    int j = i - 1;
    while (j >= 0 && xs[ord[j]] > xs[v]) {
      ord[j + 1] = ord[j];
      --j;
    }
    ord[j + 1] = v;
  }
  double w[4];
  for (int i = 0; i < nl; ++i) {
    double wi = -INFINITY;
    if (i > 0) wi = fmax(wi, xs[ord[i]] - xs[ord[i - 1]]);
    if (i + 1 < nl) wi = fmax(wi, xs[ord[i + 1]] - xs[ord[i]]);
    w[ord[i]] = wi;
  }
  if (!(M_all - m_all > mdt)) {  // every choice of hits is within the window
    for (int i = 0; i < nl; ++i) G.mark(g[i], w[i]);
    return;
  }
  double lmax = -INFINITY, umin = INFINITY;  // any choice spans at least max(earliest) - min(latest)
  for (int i = 0; i < nl; ++i) {
    lmax = fmax(lmax, G.tmin(g[i]));
    umin = fmin(umin, G.tmax(g[i]));
  }
  if (lmax - umin > mdt) return;  // no choice of hits is within the window
  for (int i = 0; i < nl; ++i)
    for (int k = G.h0(g[i]); k < G.h1(g[i]); ++k)
      if (tup_time_ok(G, g, nl, i, G.ht(k), mdt)) {
        gmax_double(&wq[k], w[i]);
        kept[k] = 1;
      }
}

template <class A>
__device__ __forceinline__ double tup_slope(const A &G, int ga, int gb) {
  return (G.x(gb) - G.x(ga)) / (G.y(gb) - G.y(ga));
}

// every tuple of the needed layer sets, split over the block by its first two groups and pruned by the widest slope
// range, slope difference and time window of the cluster (a pruned tuple fails its own, narrower cuts too; any choice
// of hits spans at least the largest earliest time minus the smallest latest time of its groups)
template <class A>
__device__ void tup_search(const A &G, const int lgo[5], int nmin, int nmax, double lo, double hi, double msd,
                           double mdtw, double *wq, uint8_t *kept) {
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int Lc[4];
    for (int i = 0; i < nl; ++i) Lc[i] = i;
    for (;;) {
      // one work item per choice of the first three groups (two for a 2-layer subset), spread over the block
      const int n1 = lgo[Lc[1] + 1] - lgo[Lc[1]], n2 = nl >= 3 ? lgo[Lc[2] + 1] - lgo[Lc[2]] : 1;
      const int P = (lgo[Lc[0] + 1] - lgo[Lc[0]]) * n1 * n2;
      bool empty = false;
      for (int i = 2; i < nl; ++i) empty |= lgo[Lc[i] + 1] == lgo[Lc[i]];
      for (int ci = empty ? P : (int)threadIdx.x; ci < P; ci += BT) {
        int g[4];
        double sl[3];
        const int rest = ci % (n1 * n2);
        g[0] = lgo[Lc[0]] + ci / (n1 * n2);
        g[1] = lgo[Lc[1]] + rest / n2;
        sl[0] = tup_slope(G, g[0], g[1]);
        if (sl[0] < lo || sl[0] > hi) continue;
        // time bound of the partial tuple: max of the groups' earliest minus min of their latest times
        const double l01 = fmax(G.tmin(g[0]), G.tmin(g[1])), u01 = fmin(G.tmax(g[0]), G.tmax(g[1]));
        if (l01 - u01 > mdtw) continue;
        if (nl == 2) {
          tup_leaf(G, g, 2, sl, wq, kept);
          continue;
        }
        g[2] = lgo[Lc[2]] + rest % n2;
        sl[1] = tup_slope(G, g[1], g[2]);
        if (sl[1] < lo || sl[1] > hi || fabs(sl[0] - sl[1]) > msd) continue;
        const double l2 = fmax(l01, G.tmin(g[2])), u2 = fmin(u01, G.tmax(g[2]));
        if (l2 - u2 > mdtw) continue;
        if (nl == 3) {
          tup_leaf(G, g, 3, sl, wq, kept);
          continue;
        }
        for (g[3] = lgo[Lc[3]]; g[3] < lgo[Lc[3] + 1]; ++g[3]) {
          sl[2] = tup_slope(G, g[2], g[3]);
          if (sl[2] < lo || sl[2] > hi || fabs(sl[0] - sl[2]) > msd || fabs(sl[1] - sl[2]) > msd) continue;
          if (fmax(l2, G.tmin(g[3])) - fmin(u2, G.tmax(g[3])) > mdtw) continue;
          tup_leaf(G, g, 4, sl, wq, kept);
        }
      }
      int k = nl - 1;
      while (k >= 0 && Lc[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++Lc[k];
      for (int i = k + 1; i < nl; ++i) Lc[i] = Lc[i - 1] + 1;
    }
  }
}

// checkCoincidence (reference reco/coinc.py:184-207) of a cluster mem[0..n) by the whole block. Writes its hits to
// byl[0..n) sorted by (layer, bar, time, pulse) with wq / kept per position. Returns 2 when the big-cluster rule keeps
// every member with its initial distance, else 1. Finite times: a search over tuples of bar groups (staged in LDS
// when the cluster fits); otherwise every hit combination, as the reference enumerates them.
__device__ __noinline__ int block_check(BlockLds &L, const crv_params &p, const Tables &t, const BS &S,
                                        const int32_t *mem, int n, int32_t *byl, double *wq, uint8_t *kept,
                                        int32_t *gs, int32_t *gid, double *gw, uint8_t *gk) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    L.flag = 0;
    for (int l = 0; l < 8; ++l) L.ishare[l] = 0;
    L.ishare[5] = 1 << 30;
    L.ishare[6] = -1;
  }
  __syncthreads();
  double lo = INFINITY, hi = -INFINITY, msd = -INFINITY, mdtw = -INFINITY;
  int nmin = 1 << 30, nmax = -1, lmask = 0;
  for (int i = tid; i < n; i += BT) {
    const int h = mem[i];
    if (!isfinite(S.t[h])) L.flag = 1;
    const Cut c = cut_of(t, S.sec[h]);
    lo = fmin(lo, c.min_slope);
    hi = fmax(hi, c.max_slope);
    msd = fmax(msd, c.max_slope_diff);
    mdtw = fmax(mdtw, c.max_dt);
    nmin = min(nmin, (int)c.coincidence_layers);
    nmax = max(nmax, (int)c.coincidence_layers);
    lmask |= 1 << S.layer[h];
  }
  atomicMin(&L.ishare[5], nmin);
  atomicMax(&L.ishare[6], nmax);
  atomicOr(&L.ishare[7], lmask);
  lo = block_min(L, lo);
  hi = block_max(L, hi);
  msd = block_max(L, msd);
  mdtw = block_max(L, mdtw);
  const bool fin = !L.flag;
  nmin = L.ishare[5];
  nmax = L.ishare[6];
  lmask = L.ishare[7];
  __syncthreads();
  if (n > p.big_cluster_threshold && __popc(lmask) >= nmin) return 2;
  // rank sort by (layer, bar, time, pulse) (time left out when a time is not finite): a total order
  const bool lm = n <= TN;
  TpLds &T = L.u.tp;
  if (lm) {
    for (int i = tid; i < n; i += BT) {
      const int h = mem[i];
      T.key[i] = S.layer[h] * t.nbars + S.bar[h];
      T.ts[i] = fin ? S.t[h] : 0.0;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += BT) {
    const int h = mem[i];
    const int ka = lm ? T.key[i] : S.layer[h] * t.nbars + S.bar[h];
    const double ta = lm ? T.ts[i] : fin ? S.t[h] : 0.0;
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const int hj = mem[j];
      const int kb = lm ? T.key[j] : S.layer[hj] * t.nbars + S.bar[hj];
      const double tb = lm ? T.ts[j] : fin ? S.t[hj] : 0.0;
      r += kb < ka || (kb == ka && (tb < ta || (tb == ta && hj < h)));
    }
    byl[r] = h;
  }
  __syncthreads();
  for (int q = tid; q < n; q += BT) {
    const int h = byl[q];
    wq[q] = cut_of(t, S.sec[h]).final_max_distance;
    kept[q] = 0;
  }
  __syncthreads();
  if (!fin) {  // every hit combination, layer by layer
    int lay[5], cmin, cmax;
    cluster_layout(t, S, byl, n, lay, cmin, cmax);
    combos(t, S, byl, wq, kept, lay, cmin, cmax, tid, BT, true);
    __syncthreads();
    return 1;
  }
  // groups: runs of one bar
  int ng = 0;
  for (int c = 0; c < n; c += BT) {
    const int q = c + tid;
    const bool f = q < n && (q == 0 || S.bar[byl[q]] != S.bar[byl[q - 1]]);
    int tot;
    const int r = bscan(L, f, tot);
    if (q < n) {
      const int g = ng + r + (f ? 0 : -1);
      gid[q] = g;
      if (f) {
        gs[g] = q;
        atomicAdd(&L.ishare[S.layer[byl[q]]], 1);
      }
    }
    ng += tot;
  }
  __syncthreads();
  if (tid == 0) gs[ng] = n;
  int lgo[5];
  lgo[0] = 0;
  for (int l = 0; l < 4; ++l) lgo[l + 1] = lgo[l] + L.ishare[l];
  __syncthreads();
  const bool lg = lm && ng <= TG;
  if (lg) {
    for (int q = tid; q < n; q += BT) T.ts[q] = S.t[byl[q]];
    for (int g = tid; g <= ng; g += BT) T.gs[g] = g < ng ? gs[g] : n;
    __syncthreads();
    for (int g = tid; g < ng; g += BT) {
      const int h = byl[T.gs[g]];
      const Cut c = cut_of(t, S.sec[h]);
      T.gx[g] = S.x[h];
      T.gy[g] = S.y[h];
      T.gtmin[g] = T.ts[T.gs[g]];
      T.gtmax[g] = T.ts[T.gs[g + 1] - 1];
      T.gcl[g] = (int8_t)max(-128, min(127, (int)c.coincidence_layers));
      T.gmdt[g] = c.max_dt;
      T.glo[g] = c.min_slope;
      T.ghi[g] = c.max_slope;
      T.gmsd[g] = c.max_slope_diff;
      T.gw[g] = -INFINITY;
      T.gk[g] = 0;
    }
    __syncthreads();
    tup_search(TupL{T}, lgo, nmin, nmax, lo, hi, msd, mdtw, wq, kept);
  } else {
    for (int g = tid; g < ng; g += BT) {
      gw[g] = -INFINITY;
      gk[g] = 0;
    }
    __syncthreads();
    tup_search(TupG{S, t, byl, gs, gw, gk}, lgo, nmin, nmax, lo, hi, msd, mdtw, wq, kept);
  }
  __syncthreads();
  for (int q = tid; q < n; q += BT) {
    const int g = gid[q];
    if (lg ? T.gk[g] : gk[g]) {
      kept[q] = 1;
      wq[q] = fmax(wq[q], lg ? T.gw[g] : gw[g]);
    }
  }
  __syncthreads();
  return 1;
}

struct BlockAcc {
  const BS &S;
  __device__ double pe(int h) const { return S.pe[h]; }
  __device__ double tm(int h) const { return S.t[h]; }
  __device__ int bar(int h) const { return S.bar[h]; }
  __device__ int sipm(int h) const { return S.sipm[h]; }
  __device__ int sec(int h) const { return S.sec[h]; }
  __device__ double x(int h) const { return S.x[h]; }
  __device__ double y(int h) const { return S.y[h]; }
  __device__ double bp(const Tables &t, int h, int ax) const { return t.bpos[S.bar[h] * 3 + ax]; }
  __device__ double mincp(const Tables &t, int h) const { return cut_of(t, S.sec[h]).min_cluster_pes; }
};

__global__ __launch_bounds__(BT) void k_coinc_block(crv_params p, Tables t, int64_t nq, const int64_t *__restrict__ eoff,
                                                    const int64_t *__restrict__ evk, const double *__restrict__ ptime,
                                                    const double *__restrict__ ppe, Ev e, BS s0,
                                                    int32_t *__restrict__ ncl, int32_t *__restrict__ nmem,
                                                    ClSlots slots, int32_t *__restrict__ cl_moff,
                                                    int32_t *__restrict__ mem, int64_t min_n,
                                                    const int32_t *__restrict__ qorder,
                                                    const uint32_t *__restrict__ qkey, int *__restrict__ qctr,
                                                    int *__restrict__ flags) {
  __shared__ BlockLds L;
  const int tid = threadIdx.x;
  // a work queue of the nq events of at least min_n pulses in decreasing pulse count (the largest first, so that no
  // block is left with a large event at the end)
  for (;;) {
    if (tid == 0) L.qpos = atomicAdd(qctr, 1);
    __syncthreads();
    const int64_t k = L.qpos;
    __syncthreads();  // every thread has read the position before thread 0 takes the next
    if (k >= nq || (int64_t)qkey[k] < min_n) break;
    const int64_t g = qorder[k];
    const int64_t a = eoff[g], n64 = eoff[g + 1] - a;
    const int n = (int)n64;
    const int64_t evkey = evk[g];
    const int64_t sb = a + g;
    BS S = s0;
    {
      double **dp[] = {&S.t, &S.pe, &S.x, &S.y, &S.maxd, &S.wm, &S.cmaxd, &S.wq, &S.cdt, &S.px, &S.pt, &S.pmd};
      for (auto q : dp) *q += sb;
      int32_t **ip[] = {&S.bar, &S.sec, &S.cit, &S.hl, &S.filt, &S.items, &S.off, &S.coinc, &S.byl, &S.cl2, &S.off2,
                        &S.bsort, &S.tsort, &S.label, &S.pass, &S.chk, &S.sk, &S.sc, &S.clmc, &S.tmp, &S.clid,
                        &S.tbin};
      for (auto q : ip) *q += sb;
      uint8_t **up[] = {&S.layer, &S.side, &S.sipm, &S.seq0, &S.kept, &S.fpass, &S.memf, &S.pside};
      for (auto q : up) *q += sb;
    }
    int32_t *cl_moff_e = cl_moff + sb;
    int32_t *mem_e = mem + a;
    int nclus = 0, nm = 0;
    for (int k = tid; k < n; k += BT) {
      S.bar[k] = e.bar[a + k];
      S.sipm[k] = (uint8_t)(e.sipm[a + k] & 0xFF);
      S.seq0[k] = e.seq[a + k] == 0;
    }
    __syncthreads();
    for (int ti = 0; ti < t.ntypes; ++ti) {
      const int type = t.type_id[ti];
      const double P0 = t.tp0[ti], P1 = t.tp1[ti], half = t.thalf[ti], tmaxdt = t.tmaxdt[ti];
      const bool both = t.tside0[ti] && t.tside1[ti];
      // hits of this sector type, in pulse order
      int nh = 0;
      for (int c = 0; c < n; c += BT) {
        const int k = c + tid;
        bool hit = false;
        if (k < n && !(p.ignore_secondary_pulses && !S.seq0[k])) {
          const int bar = e.bar[a + k], sp = e.sipm[a + k];
          if (bar < 0 || bar >= t.nbars || sp < 0 || sp > 3 || bar * 4 + 3 >= t.n_channels) {
            atomicAdd(&flags[F_BADCHAN], 1);
          } else {
            const BarInfo bi = t.binfo[bar];
            if (bi.type == type) {
              double tt = ptime[a + k];
              if (sp % 2 == 0) tt += (bi.ps0 - P0) / p.fiber_signal_speed;
              else tt += (P1 - bi.ps1) / p.fiber_signal_speed;
              double pe = ppe[a + k];
              const int other = bar * 4 + sp + (sp < 2 ? 2 : -2);
              if (t.status[other] & p.compensate_mask) pe *= 2.0;
              S.t[k] = tt;
              S.pe[k] = pe;
              S.maxd[k] = bi.imd;
              S.x[k] = bi.x;
              S.y[k] = bi.y;
              S.sec[k] = bi.sec;
              S.layer[k] = (uint8_t)bi.layer;
              S.cit[k] = bi.cit;
              S.side[k] = (uint8_t)(sp % 2);
              hit = true;
            }
          }
        }
        int tot;
        const int r = bscan(L, hit, tot);
        if (hit) S.hl[nh + r] = k;
        nh += tot;
      }
      __syncthreads();
      if (nh == 0) continue;
      // filterHits (reco/coinc.py:154-182): buckets of (side, layer, counter column) in hit order
      const int nbk = 8 * t.ncit;
      for (int b = tid; b < nbk + 1; b += BT) L.u.hist[b] = 0;
      __syncthreads();
      for (int ii = tid; ii < nh; ii += BT) {
        const int h = S.hl[ii];
        atomicAdd(&L.u.hist[(S.side[h] * 4 + S.layer[h]) * t.ncit + S.cit[h]], 1);
      }
      __syncthreads();
      bscan_hist(L, nbk + 1);
      for (int ii = tid; ii < nh; ii += BT) {
        const int h = S.hl[ii];
        S.bsort[atomicAdd(&L.u.hist[(S.side[h] * 4 + S.layer[h]) * t.ncit + S.cit[h]], 1)] = h;
      }
      __syncthreads();  // hist[b] is now the end of bucket b; order each bucket by pulse (= hit order)
      for (int b = tid; b < nbk; b += BT) {
        const int b0 = b ? L.u.hist[b - 1] : 0, b1 = L.u.hist[b];
        for (int i = b0 + 1; i < b1; ++i) {
          const int v = S.bsort[i];
          int j = i - 1;
          while (j >= b0 && S.bsort[j] > v) {
            S.bsort[j + 1] = S.bsort[j];
            --j;
          }
          S.bsort[j + 1] = v;
        }
      }
      __syncthreads();
      for (int ii = tid; ii < nh; ii += BT) {
        const int h = S.hl[ii];
        const Cut ch = cut_of(t, S.sec[h]);
        const int row = (S.side[h] * 4 + S.layer[h]) * t.ncit, cnh = S.cit[h];
        double sums[3] = {0.0, 0.0, 0.0};  // offset -1, 0, +1
        for (int d = -1; d <= 1; ++d) {
          if (cnh + d < 0 || cnh + d >= t.ncit) continue;
          const int b = row + cnh + d;
          for (int k = b ? L.u.hist[b - 1] : 0; k < L.u.hist[b]; ++k) {
            const int q = S.bsort[k];
            if (fabs(S.t[q] - S.t[h]) > ch.max_dt_adjacent) continue;
            sums[d + 1] += S.pe[q];
          }
        }
        S.fpass[ii] = sums[1] + sums[0] >= ch.pe_threshold || sums[1] + sums[2] >= ch.pe_threshold;
      }
      __syncthreads();
      // per side: passing hits in hit order, their clusters and the coincidence check of each
      int nco = 0, nfilt = 0;
      for (int side = 0; side < 2; ++side) {
        const int f0 = nfilt;
        for (int c = 0; c < nh; c += BT) {
          const int ii = c + tid;
          const bool keep = ii < nh && S.side[S.hl[ii]] == side && S.fpass[ii];
          int tot;
          const int r = bscan(L, keep, tot);
          if (keep) S.filt[nfilt + r] = S.hl[ii];
          nfilt += tot;
        }
        __syncthreads();
        // the clusters go to cl2 / off2 as scratch here (the second clustering reuses them later)
        const int ncs = block_clusters(L, S, S.filt + f0, nfilt - f0, S.maxd, tmaxdt, false, 0.0, S.cl2, S.off2);
        // coincidence check: a small cluster's by-layer lists are built by one thread and its combinations spread
        // over the block with the other small clusters' (one flat index space: S.sc holds each cluster's first
        // combination, S.tmp its packed layer offsets); the block checks the others cluster by cluster
        int nbig = 0, nflat = 0;
        for (int c0 = 0; c0 < ncs; c0 += BT) {
          const int c = c0 + tid;
          bool big = false;
          int pc = 0;
          if (c < ncs) {
            const int o0 = S.off2[c], mn = S.off2[c + 1] - o0;
            S.clmc[c] = 1;  // 1: checked by the block
            if (mn <= SMALL_CLUSTER) {
              int lay[5], nmin, nmax;
              bool allk;
              const int64_t P = thread_check(p, t, S, S.cl2 + o0, mn, S.byl + o0, S.wq + o0, S.kept + o0, lay, nmin,
                                             nmax, allk);
              if (allk) S.clmc[c] = -1;  // -1: every member kept with its initial distance
              else if (P <= BIG_COMBOS) {
                S.clmc[c] = 0;  // 0: checked in the flat combination space
                pc = (int)P;
                S.tmp[c] = lay[1] | lay[2] << 6 | lay[3] << 12 | lay[4] << 18 | nmin << 24 | nmax << 27;
              }
            }
            big = S.clmc[c] == 1;
          }
          int tot;
          const int r = bscan(L, big, tot);
          if (big) S.sk[nbig + r] = c;
          nbig += tot;
          const int rp = bscan(L, pc, tot);
          if (c < ncs) S.sc[c] = nflat + rp;
          nflat += tot;
        }
        __syncthreads();
        for (int f = tid; f < nflat; f += BT) {
          int lo = 0, hi = ncs - 1;  // the last cluster whose first combination is <= f (it has combinations)
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.sc[mid] <= f) lo = mid;
            else hi = mid - 1;
          }
          const int c = lo, o0 = S.off2[c], pk = S.tmp[c];
          const int lay[5] = {0, pk & 63, (pk >> 6) & 63, (pk >> 12) & 63, (pk >> 18) & 63};
          int idx[4];
          const int nl = combo_at(f - S.sc[c], lay, (pk >> 24) & 7, (pk >> 27) & 7, idx);
          if (nl) combo_eval(t, S, S.byl + o0, S.wq + o0, S.kept + o0, idx, nl, true);
        }
        __syncthreads();
        for (int bi = 0; bi < nbig; ++bi) {
          const int c = S.sk[bi], o0 = S.off2[c], mn = S.off2[c + 1] - o0;
          const int rc = block_check(L, p, t, S, S.cl2 + o0, mn, S.byl + o0, S.wq + o0, S.kept + o0, S.label + o0,
                                     S.tbin + o0, S.cdt + o0, S.memf + o0);
          if (rc == 2 && tid == 0) S.clmc[c] = -1;
          __syncthreads();
        }
        // kept hits per cluster in pulse order, appended in cluster order with their widened distance; the hits of
        // the small clusters are ranked one per thread over all of them (S.tmp: each cluster's first output slot,
        // S.sc: its first hit in that flat list)
        int nfh = 0;
        for (int c0 = 0; c0 < ncs; c0 += BT) {
          const int c = c0 + tid;
          int cnt = 0, fh = 0;
          if (c < ncs) {
            const int o0 = S.off2[c], mn = S.off2[c + 1] - o0;
            if (S.clmc[c] < 0) cnt = mn;
            else
              for (int q = 0; q < mn; ++q) cnt += S.kept[o0 + q];
            if (S.clmc[c] == 0 && cnt) fh = mn;
          }
          int tot;
          const int r = bscan(L, cnt, tot);
          if (c < ncs && cnt) {
            const int o0 = S.off2[c], mn = S.off2[c + 1] - o0, w0 = nco + r;
            if (S.clmc[c] < 0) {
              for (int i = 0; i < mn; ++i) {
                S.coinc[w0 + i] = S.cl2[o0 + i];
                S.cmaxd[w0 + i] = S.maxd[S.cl2[o0 + i]];
              }
            } else {
              S.tmp[c] = w0;
            }
          }
          nco += tot;
          const int rf = bscan(L, fh, tot);
          if (c < ncs) S.sc[c] = nfh + rf;
          nfh += tot;
        }
        __syncthreads();
        for (int f = tid; f < nfh; f += BT) {  // rank by pulse among the cluster's kept hits
          int lo = 0, hi = ncs - 1;  // the last cluster whose first flat hit is <= f (it has hits in the list)
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.sc[mid] <= f) lo = mid;
            else hi = mid - 1;
          }
          const int c = lo, o0 = S.off2[c], mn = S.off2[c + 1] - o0, q = f - S.sc[c];
          if (!S.kept[o0 + q]) continue;
          const int h = S.byl[o0 + q];
          int rk = 0;
          for (int q2 = 0; q2 < mn; ++q2) rk += S.kept[o0 + q2] && S.byl[o0 + q2] < h;
          S.coinc[S.tmp[c] + rk] = h;
          S.cmaxd[S.tmp[c] + rk] = S.wq[o0 + q];
        }
        __syncthreads();
        for (int bi = 0; bi < nbig; ++bi) {  // kept hits of the block-checked clusters: rank by pulse
          const int c = S.sk[bi];
          if (S.clmc[c] != 1) continue;
          const int o0 = S.off2[c], mn = S.off2[c + 1] - o0, w0 = S.tmp[c];
          const bool lm = mn <= BH;
          if (lm) {
            for (int q = tid; q < mn; q += BT) L.u.hist[q] = S.kept[o0 + q] ? S.byl[o0 + q] : INT_MAX;
            __syncthreads();
          }
          for (int q = tid; q < mn; q += BT) {
            if (!S.kept[o0 + q]) continue;
            const int h = S.byl[o0 + q];
            int r = 0;
            for (int q2 = 0; q2 < mn; ++q2)
              r += (lm ? L.u.hist[q2] : S.kept[o0 + q2] ? S.byl[o0 + q2] : INT_MAX) < h;
            S.coinc[w0 + r] = h;
            S.cmaxd[w0 + r] = S.wq[o0 + q];
          }
          __syncthreads();
        }
      }
      const double dt_opp = 2.0 * half / p.fiber_signal_speed;
      for (int i = tid; i < nco; i += BT) S.wm[S.coinc[i]] = S.cmaxd[i];
      __syncthreads();
      const int ncl2 = block_clusters(L, S, S.coinc, nco, S.wm, tmaxdt, both, dt_opp, S.cl2, S.off2);
      const int nitems = ncl2 ? S.off2[ncl2] : 0;
      // attachment of opposite-side pulses (reco/coinc_cpu.py:11-51), candidates looked up by bar
      int nchosen = 0;
      if (both && p.add_opposite_side_pulses && nitems > 0 && nfilt > 0) {
        for (int b = tid; b < t.nbars + 1; b += BT) L.u.hist[b] = 0;
        for (int k = tid; k < n; k += BT) S.memf[k] = 0;
        __syncthreads();
        for (int c = tid; c < ncl2; c += BT)
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) {
            S.clid[m] = c;
            S.memf[S.cl2[m]] = 1;
            atomicAdd(&L.u.hist[S.bar[S.cl2[m]]], 1);
          }
        __syncthreads();
        bscan_hist(L, t.nbars + 1);
        for (int m = tid; m < nitems; m += BT) S.bsort[atomicAdd(&L.u.hist[S.bar[S.cl2[m]]], 1)] = m;
        __syncthreads();  // hist[bar] = end of the bar's items
        const double mdt = tmaxdt + dt_opp;
        for (int ci = tid; ci < nfilt; ci += BT) {
          const int h = S.filt[ci];
          int best_c = -1;
          double best_dt = 0.0;
          if (S.seq0[h] && !S.memf[h]) {
            const int hb = S.bar[h], hs = S.side[h];
            const int k0 = hb ? L.u.hist[hb - 1] : 0, k1 = L.u.hist[hb];
            for (int k = k0; k < k1; ++k) {
              const int m = S.bsort[k], an = S.cl2[m], c = S.clid[m];
              if (!S.seq0[an] || hs != 1 - S.side[an]) continue;
              const double dt = fabs(S.t[h] - S.t[an]);
              if (!(dt < mdt)) continue;
              bool occ = false;
              for (int kk = k0; kk < k1 && !occ; ++kk)
                occ = S.clid[S.bsort[kk]] == c && S.side[S.cl2[S.bsort[kk]]] == hs;
              if (occ) continue;
              if (best_c < 0 || dt < best_dt || (dt == best_dt && c < best_c)) {
                best_c = c;
                best_dt = dt;
              }
            }
          }
          S.chk[ci] = best_c;
          S.cdt[ci] = best_dt;
        }
        __syncthreads();
        for (int c0 = 0; c0 < nfilt; c0 += BT) {  // the chosen candidates, in candidate order (S.tmp)
          const int ci = c0 + tid;
          const bool chosen = ci < nfilt && S.chk[ci] >= 0;
          int tot;
          const int r = bscan(L, chosen, tot);
          if (chosen) S.tmp[nchosen + r] = ci;
          nchosen += tot;
        }
        __syncthreads();
        for (int u = tid; u < nchosen; u += BT) {  // rank in (cluster, bar, sipm, dt, candidate) order
          const int ci = S.tmp[u];
          const int k = S.chk[ci], h = S.filt[ci], b = S.bar[h], sp = S.sipm[h];
          const double dt = S.cdt[ci];
          int rank = 0;
          for (int v = 0; v < nchosen; ++v) {
            const int cj = S.tmp[v];
            if (cj == ci) continue;
            const int kj = S.chk[cj], hj = S.filt[cj], bj = S.bar[hj], sj = S.sipm[hj];
            const double dj = S.cdt[cj];
            rank += kj < k || (kj == k && (bj < b || (bj == b && (sj < sp || (sj == sp && (dj < dt ||
                    (dj == dt && cj < ci)))))));
          }
          S.sk[rank] = k;
          S.sc[rank] = ci;
        }
        __syncthreads();
      }
      // members (items, then attachments one per (bar, sipm)) and the min-PE test per cluster; offsets; moments
      for (int c = tid; c < ncl2; c += BT) {
        int mc = S.off2[c + 1] - S.off2[c];
        double tot = 0.0, mincp = INFINITY;
        for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) {
          const int h = S.cl2[m];
          tot = tot + S.pe[h];
          mincp = fmin(mincp, cut_of(t, S.sec[h]).min_cluster_pes);
        }
        int q0 = 0, q1 = nchosen;  // attachments of cluster c: a range of the sorted list
        while (q0 < q1) {
          const int mid = (q0 + q1) >> 1;
          if (S.sk[mid] < c) q0 = mid + 1;
          else q1 = mid;
        }
        int lastb = -1, lasts = -1;
        for (int q = q0; q < nchosen && S.sk[q] == c; ++q) {
          const int h = S.filt[S.sc[q]];
          if (S.bar[h] == lastb && S.sipm[h] == lasts) continue;
          lastb = S.bar[h];
          lasts = S.sipm[h];
          ++mc;
          tot = tot + S.pe[h];
          mincp = fmin(mincp, cut_of(t, S.sec[h]).min_cluster_pes);
        }
        S.clmc[c] = tot < mincp ? 0 : mc;
      }
      __syncthreads();
      for (int c0 = 0; c0 < ncl2; c0 += BT) {
        const int c = c0 + tid;
        const int mc = c < ncl2 ? S.clmc[c] : 0;
        int mtot, ktot;
        const int mr = bscan(L, mc, mtot);
        const int kr = bscan(L, mc > 0, ktot);
        if (mc > 0) {
          const int mbase = nm + mr, cslot = nclus + kr;
          int32_t *mm = mem_e + mbase;
          int w = 0;
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) mm[w++] = S.cl2[m];
          int q0 = 0, q1 = nchosen;
          while (q0 < q1) {
            const int mid = (q0 + q1) >> 1;
            if (S.sk[mid] < c) q0 = mid + 1;
            else q1 = mid;
          }
          int lastb = -1, lasts = -1;
          for (int q = q0; q < nchosen && S.sk[q] == c; ++q) {
            const int h = S.filt[S.sc[q]];
            if (S.bar[h] == lastb && S.sipm[h] == lasts) continue;
            lastb = S.bar[h];
            lasts = S.sipm[h];
            mm[w++] = h;
          }
          cluster_emit(p, t, BlockAcc{S}, mm, mc, type, P0, P1, evkey, a + cslot, slots);
          cl_moff_e[cslot] = mbase;
        }
        nm += mtot;
        nclus += ktot;
      }
      __syncthreads();
    }
    if (tid == 0) {
      cl_moff_e[nclus] = nm;
      ncl[g] = nclus;
      nmem[g] = nm;
    }
    __syncthreads();
  }
}

// the block finder's events: each event of at least min_n pulses appended (in no particular order) with its pulse
// count, at most 65535, as the queue's sort key
// and the wavefront finders' events: class c (0: <= 64 pulses, 1: <= 128, 2: <= 256) listed at lists + c * nev, in no
// particular order, their numbers in count[1 + c] (one atomic per wavefront and class)
__global__ void k_block_keys(int64_t nev, const int64_t *__restrict__ eoff, int64_t min_n, uint32_t *__restrict__ key,
                             int32_t *__restrict__ idx, int *__restrict__ count, int32_t *__restrict__ lists,
                             const int32_t *__restrict__ nev_dev = nullptr) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nev || (nev_dev && g >= *nev_dev)) return;
  const int64_t n = eoff[g + 1] - eoff[g];
  const int cls = n <= 64 ? 0 : n <= 128 ? 1 : n < min_n ? 2 : 3;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const uint64_t bm = __ballot(cls == c);
    if (!bm) continue;
    const int leader = __ffsll((long long)bm) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&count[1 + c], __popcll(bm));
    base = __shfl(base, leader);
    if (cls == c) lists[c * nev + base + __popcll(bm & lanes_below())] = (int32_t)g;
  }
  if (cls < 3) return;
  const int k = atomicAdd(count, 1);
  key[k] = (uint32_t)(n < 65535 ? n : 65535);
  idx[k] = (int32_t)g;
}

__global__ void k_coinc_out(int64_t nev, const int64_t *__restrict__ eoff, const int64_t *__restrict__ clo,
                            const int64_t *__restrict__ mo, const int32_t *__restrict__ ncl,
                            const int64_t *__restrict__ cl_ev, const int32_t *__restrict__ cl_type,
                            const int32_t *__restrict__ cl_poscalc, const int32_t *__restrict__ cl_sidehits,
                            const double *__restrict__ cl_d, const int32_t *__restrict__ cl_moff,
                            const int32_t *__restrict__ mem, crv_clusters o,
                            const int32_t *__restrict__ nev_dev = nullptr) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nev_dev) {  // fixed-capacity calls: the events of the call, and the member offsets' end when there are none
    nev = min(nev, (int64_t)*nev_dev);
    if (nev == 0 && g == 0) o.cl_mem_off[0] = 0;
  }
  if (g >= nev) return;
  int64_t a = eoff[g], c0 = clo[g], m0 = mo[g];
  int nc = ncl[g];
  for (int k = 0; k < nc; ++k) {
    int64_t src = a + k, dst = c0 + k;
    o.cl_event[dst] = cl_ev[src];
    o.cl_type[dst] = cl_type[src];
    o.cl_poscalc[dst] = cl_poscalc[src];
    o.cl_sidehits[dst * 2] = cl_sidehits[src * 2];
    o.cl_sidehits[dst * 2 + 1] = cl_sidehits[src * 2 + 1];
    const double *d = cl_d + src * 12;
    o.cl_pe[dst] = d[0];
    o.cl_start[dst] = d[1];
    o.cl_end[dst] = d[2];
    o.cl_slope[dst] = d[3];
    o.cl_avgtime[dst] = d[4];
    for (int x = 0; x < 3; ++x) o.cl_pos[dst * 3 + x] = d[5 + x];
    o.cl_sidepe[dst * 2] = d[8];
    o.cl_sidepe[dst * 2 + 1] = d[9];
    o.cl_sidetime[dst * 2] = d[10];
    o.cl_sidetime[dst * 2 + 1] = d[11];
    int32_t mb = cl_moff[a + g + k], me = cl_moff[a + g + k + 1];
    o.cl_mem_off[dst] = m0 + mb;
    for (int32_t m = mb; m < me; ++m) o.mem_idx[m0 + m] = mem[a + m];
  }
  if (g == nev - 1) o.cl_mem_off[c0 + nc] = m0 + cl_moff[a + g + nc];
}

}  // namespace

int clusters_stage(Ctx *c, const int64_t *ev, const int32_t *bar, const int32_t *sipm, const double *time,
                   const double *pe, const int64_t *seq, int64_t n, crv_clusters *out, hipStream_t s) {
  if (!c->t.cuts) return set_error(c, CRV_E_STATE, "clusters: geometry not uploaded");
  if (n == 0) {  // no clusters: the member offsets are the single 0
    out->n = 0;
    out->n_mem = 0;
    if (out->capacity < 0) {
      out->cl_mem_off = (int64_t *)ws(c, 352, 8);
      if (!out->cl_mem_off) return ws_fail(c);
      out->capacity = 0;
      out->mem_capacity = 0;
    }
    if (out->cl_mem_off) {
      CRV_HIP(c, hipMemsetAsync(out->cl_mem_off, 0, 8, s));
      CRV_HIP(c, hipStreamSynchronize(s));
    }
    return CRV_OK;
  }
  int64_t *evk = (int64_t *)ws(c, 90, n * 8);
  int32_t *cnt = (int32_t *)ws(c, 91, (n + 1) * 4);
  int32_t *nruns = (int32_t *)ws(c, 92, 8);
  int64_t *eoff = (int64_t *)ws(c, 93, (n + 1) * 8);
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, ev, evk, cnt, nruns, (int)n, s));
  void *tmp = ws(c, 94, tb);
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp, tb, ev, evk, cnt, nruns, (int)n, s));
  int32_t hn = 0;
  CRV_HIP(c, hipMemcpyAsync(&hn, nruns, 4, hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  const int64_t nev = hn;
  int rc = scan_i32_to_i64(c, cnt, eoff, nev, 95, s);
  if (rc) return rc;
  const int64_t ns = n + nev + 2;  // event-local scratch: every event owns n + 1 entries
  int32_t *ncl = (int32_t *)ws(c, 113, (nev + 1) * 4);
  int32_t *nmem = (int32_t *)ws(c, 114, (nev + 1) * 4);
  int64_t *cl_ev = (int64_t *)ws(c, 115, (n + 1) * 8);
  int32_t *cl_type = (int32_t *)ws(c, 116, (n + 1) * 4);
  int32_t *cl_pc = (int32_t *)ws(c, 117, (n + 1) * 4);
  int32_t *cl_sh = (int32_t *)ws(c, 118, (n + 1) * 8);
  double *cl_d = (double *)ws(c, 119, (n + 1) * 12 * 8);
  int32_t *cl_moff = (int32_t *)ws(c, 120, ns * 4);
  int32_t *mem = (int32_t *)ws(c, 121, (n + 1) * 4);
  int64_t *clo = (int64_t *)ws(c, 122, (nev + 1) * 8);
  int64_t *mo = (int64_t *)ws(c, 123, (nev + 1) * 8);
  if (!ncl || !nmem || !cl_ev || !cl_type || !cl_pc || !cl_sh || !cl_d || !cl_moff || !mem || !clo || !mo)
    return ws_fail(c);
  Ev e{bar, sipm, seq};
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  kmark(c, 3, 0, s);
  {  // size classes: wavefront kernels for events of <= 64, <= 128 and <= 256 pulses, a block per larger event.
    // Each event belongs to one class, so the 128-pulse class and the 256-pulse class + block finder run on two side
    // streams beside the 64-pulse class and are joined back to `s` before anything reads their outputs.
    if (c->t.nbars + 1 > BH || 8 * c->t.ncit + 1 > BH)
      return set_error(c, CRV_E_ARG, "clusters: geometry too large for the block finder's bar / column buckets");
    const ClSlots slots{cl_ev, cl_type, cl_pc, cl_sh, cl_d};
    BS b{};
    double **bd[] = {&b.t, &b.pe, &b.x, &b.y, &b.maxd, &b.wm, &b.cmaxd, &b.wq, &b.cdt, &b.px, &b.pt, &b.pmd};
    for (int i = 0; i < 12; ++i) *bd[i] = (double *)ws(c, i < 9 ? 130 + i : 171 + i, ns * 8);
    int32_t **bi[] = {&b.bar, &b.sec, &b.cit, &b.hl, &b.filt, &b.items, &b.off, &b.coinc, &b.byl, &b.cl2, &b.off2,
                      &b.bsort, &b.tsort, &b.label, &b.pass, &b.chk, &b.sk, &b.sc, &b.clmc, &b.tmp, &b.clid, &b.tbin};
    for (int i = 0; i < 22; ++i) *bi[i] = (int32_t *)ws(c, 140 + i, ns * 4);
    uint8_t **bu[] = {&b.layer, &b.side, &b.sipm, &b.seq0, &b.kept, &b.fpass, &b.memf, &b.pside};
    for (int i = 0; i < 8; ++i) *bu[i] = (uint8_t *)ws(c, i < 7 ? 170 + i : 183, ns + 8);
    for (auto q : bd) if (!*q) return ws_fail(c);
    for (auto q : bi) if (!*q) return ws_fail(c);
    for (auto q : bu) if (!*q) return ws_fail(c);
    // the block finder's queue: the events of more than 256 pulses, compacted here (their count read back), then
    // sorted by decreasing pulse count (16-bit keys) on the block finder's side stream
    uint32_t *qk = (uint32_t *)ws(c, 186, nev * 4 + 4), *qk2 = (uint32_t *)ws(c, 187, nev * 4 + 4);
    int32_t *qi = (int32_t *)ws(c, 188, nev * 4 + 4), *qo = (int32_t *)ws(c, 189, nev * 4 + 4);
    // queue position, the number of block-finder events, the numbers of the three wavefront classes' events
    int *qctr = (int *)ws(c, 190, 64);
    int32_t *wl = (int32_t *)ws(c, 192, nev * 12 + 16);  // the wavefront classes' event lists
    if (!qk || !qk2 || !qi || !qo || !qctr || !wl) return ws_fail(c);
    CRV_HIP(c, hipMemsetAsync(qctr, 0, 20, s));
    k_block_keys<<<nblocks(nev, 256), 256, 0, s>>>(nev, eoff, (int64_t)257, qk, qi, qctr + 1, wl);
    CRV_HIP(c, hipGetLastError());
    int hc[4] = {0, 0, 0, 0};
    CRV_HIP(c, hipMemcpyAsync(hc, qctr + 1, sizeof(hc), hipMemcpyDeviceToHost, s));
    CRV_HIP(c, hipStreamSynchronize(s));
    const int nq = hc[0];
    size_t qtb = 0;
    void *qtmp = nullptr;
    if (nq > 0) {
      CRV_HIP(c, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, qtb, qk, qk2, qi, qo, nq, 0, 16, c->aux[1]));
      if (!(qtmp = ws(c, 191, qtb))) return ws_fail(c);
    }
    CRV_HIP(c, hipEventRecord(c->fork[0], s));
    CRV_HIP(c, hipStreamWaitEvent(c->aux[0], c->fork[0], 0));
    CRV_HIP(c, hipStreamWaitEvent(c->aux[1], c->fork[0], 0));
    hipError_t e0 = hipSuccess, e1 = hipSuccess, e2 = hipSuccess;
    if (hc[1]) {
      k_coinc_wave<64><<<std::min(hc[1], 16384), 64, 0, s>>>(c->p, c->t, wl, qctr + 2, eoff, evk, time, pe, e, ncl,
                                                             nmem, slots, cl_moff, mem, c->flags);
      e0 = hipGetLastError();
    }
    if (hc[2]) {
      k_coinc_wave<128><<<std::min(hc[2], 16384), 64, 0, c->aux[0]>>>(c->p, c->t, wl + nev, qctr + 3, eoff, evk, time,
                                                                      pe, e, ncl, nmem, slots, cl_moff, mem,
                                                                      c->flags);
      e1 = hipGetLastError();
    }
    if (hc[3]) {
      k_coinc_wave<256><<<std::min(hc[3], 4096), 64, 0, c->aux[1]>>>(c->p, c->t, wl + 2 * nev, qctr + 4, eoff, evk,
                                                                     time, pe, e, ncl, nmem, slots, cl_moff, mem,
                                                                     c->flags);
      e2 = hipGetLastError();
    }
    hipError_t e3 = hipSuccess;
    if (nq > 0) {
      e3 = hipcub::DeviceRadixSort::SortPairsDescending(qtmp, qtb, qk, qk2, qi, qo, nq, 0, 16, c->aux[1]);
      if (e3 == hipSuccess) {
        k_coinc_block<<<std::min(nq, 2048), BT, 0, c->aux[1]>>>(c->p, c->t, (int64_t)nq, eoff, evk, time, pe, e, b, ncl,
                                                               nmem, slots, cl_moff, mem, (int64_t)257, qo, qk2, qctr,
                                                               c->flags);
        e3 = hipGetLastError();
      }
    }
    // join before any error return, so that nothing of this call is left running on a side stream
    CRV_HIP(c, hipEventRecord(c->fork[1], c->aux[0]));
    CRV_HIP(c, hipEventRecord(c->fork[2], c->aux[1]));
    CRV_HIP(c, hipStreamWaitEvent(s, c->fork[1], 0));
    CRV_HIP(c, hipStreamWaitEvent(s, c->fork[2], 0));
    CRV_HIP(c, e0);
    CRV_HIP(c, e1);
    CRV_HIP(c, e2);
    CRV_HIP(c, e3);
  }
  kmark(c, 3, 1, s);
  if ((rc = scan_i32_to_i64(c, ncl, clo, nev, 124, s))) return rc;
  if ((rc = scan_i32_to_i64(c, nmem, mo, nev, 125, s))) return rc;
  int64_t nc = 0, nmt = 0;
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  if ((rc = read_i64(c, clo + nev, &nc, s))) return rc;
  if (hf[F_BADCHAN]) return set_error(c, CRV_E_ARG, "clusters: bar or SiPM outside the uploaded geometry");
  if ((rc = read_i64(c, mo + nev, &nmt, s))) return rc;
  out->n = nc;
  out->n_mem = nmt;
  if (out->capacity < 0) {  // internal outputs of the chained run
    out->cl_event = (int64_t *)ws(c, 340, nc * 8 + 8);
    out->cl_type = (int32_t *)ws(c, 341, nc * 4 + 8);
    out->cl_poscalc = (int32_t *)ws(c, 342, nc * 4 + 8);
    out->cl_sidehits = (int32_t *)ws(c, 343, nc * 8 + 8);
    out->cl_pe = (double *)ws(c, 344, nc * 8 + 8);
    out->cl_start = (double *)ws(c, 345, nc * 8 + 8);
    out->cl_end = (double *)ws(c, 346, nc * 8 + 8);
    out->cl_slope = (double *)ws(c, 347, nc * 8 + 8);
    out->cl_avgtime = (double *)ws(c, 348, nc * 8 + 8);
    out->cl_pos = (double *)ws(c, 349, nc * 24 + 8);
    out->cl_sidepe = (double *)ws(c, 350, nc * 16 + 8);
    out->cl_sidetime = (double *)ws(c, 351, nc * 16 + 8);
    out->cl_mem_off = (int64_t *)ws(c, 352, (nc + 1) * 8);
    out->mem_idx = (int64_t *)ws(c, 353, nmt * 8 + 8);
    out->capacity = nc;
    out->mem_capacity = nmt;
  }
  if (nc > out->capacity || nmt > out->mem_capacity) return set_error(c, CRV_E_CAPACITY, "clusters: output capacity too small");
  k_coinc_out<<<nblocks(nev, 256), 256, 0, s>>>(nev, eoff, clo, mo, ncl, cl_ev, cl_type, cl_pc, cl_sh, cl_d, cl_moff,
                                                mem, *out);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipStreamSynchronize(s));
  return CRV_OK;
}


namespace {

// events of a fixed-capacity call: the runs of real event keys (the padding rows, key INT64_MAX, form one more run);
// the pulse counts past them are cleared, so the event offsets of the padding events all equal the pulse total
__global__ void k_fix_events(const int32_t *__restrict__ nruns, const int64_t *__restrict__ np, int64_t n_cap,
                             int32_t *__restrict__ cnt, int32_t *__restrict__ nev_dev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = *np;
  const int32_t nev = n == 0 ? 0 : *nruns - (n < n_cap ? 1 : 0);
  if (i == 0) *nev_dev = nev;
  if (i >= nev && i <= n_cap) cnt[i] = 0;
}

__global__ void k_zero_i32(int32_t *__restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_cluster_sizes(const int64_t *__restrict__ clo, const int64_t *__restrict__ mo, int64_t nev_cap,
                                const int *__restrict__ skip, int64_t *__restrict__ sz) {
  sz[1] = *skip ? 0 : clo[nev_cap];
  sz[2] = *skip ? 0 : mo[nev_cap];
}

}  // namespace

// clusters_stage over a fixed capacity of n_cap pulse rows, rows past the call's pulses carrying the event key
// INT64_MAX: every event of the capacity runs through the size-class finders (the padding events are empty and find
// nothing), the block finder takes its events from a queue sorted over the whole capacity (non-queue entries have
// key 0 and end it); outputs at the plan's capacity.
int clusters_fixed(Ctx *c, Plan &pl, Arena &ar, const int64_t *ev, const int32_t *bar, const int32_t *sipm,
                   const double *time, const double *pe, const int64_t *seq, hipStream_t s) {
  if (!c->t.cuts) return set_error(c, CRV_E_STATE, "clusters: geometry not uploaded");
  if (c->t.nbars + 1 > BH || 8 * c->t.ncit + 1 > BH)
    return set_error(c, CRV_E_ARG, "clusters: geometry too large for the block finder's bar / column buckets");
  const int64_t n = pl.np_cap, nev = n + 1;  // event capacity: every pulse its own event, and the padding event
  auto A = [&](size_t b) { return ar.get(b); };
  int64_t *evk = (int64_t *)A(nev * 8);
  int32_t *cnt = (int32_t *)A((nev + 1) * 4);
  int32_t *nruns = (int32_t *)A(8), *nev_dev = (int32_t *)A(8);
  int64_t *eoff = (int64_t *)A((nev + 1) * 8);
  const int64_t ns = n + nev + 2;
  int32_t *ncl = (int32_t *)A((nev + 1) * 4), *nmem = (int32_t *)A((nev + 1) * 4);
  int64_t *cl_ev = (int64_t *)A((n + 1) * 8);
  int32_t *cl_type = (int32_t *)A((n + 1) * 4), *cl_pc = (int32_t *)A((n + 1) * 4);
  int32_t *cl_sh = (int32_t *)A((n + 1) * 8);
  double *cl_d = (double *)A((n + 1) * 12 * 8);
  int32_t *cl_moff = (int32_t *)A(ns * 4), *mem = (int32_t *)A((n + 1) * 4);
  int64_t *clo = (int64_t *)A((nev + 1) * 8), *mo = (int64_t *)A((nev + 1) * 8);
  BS b{};
  double **bd[] = {&b.t, &b.pe, &b.x, &b.y, &b.maxd, &b.wm, &b.cmaxd, &b.wq, &b.cdt, &b.px, &b.pt, &b.pmd};
  for (auto q : bd) *q = (double *)A(ns * 8);
  int32_t **bi[] = {&b.bar, &b.sec, &b.cit, &b.hl, &b.filt, &b.items, &b.off, &b.coinc, &b.byl, &b.cl2, &b.off2,
                    &b.bsort, &b.tsort, &b.label, &b.pass, &b.chk, &b.sk, &b.sc, &b.clmc, &b.tmp, &b.clid, &b.tbin};
  for (auto q : bi) *q = (int32_t *)A(ns * 4);
  uint8_t **bu[] = {&b.layer, &b.side, &b.sipm, &b.seq0, &b.kept, &b.fpass, &b.memf, &b.pside};
  for (auto q : bu) *q = (uint8_t *)A(ns + 8);
  uint32_t *qk = (uint32_t *)A(nev * 4 + 4), *qk2 = (uint32_t *)A(nev * 4 + 4);
  int32_t *qi = (int32_t *)A(nev * 4 + 4), *qo = (int32_t *)A(nev * 4 + 4);
  int *qctr = (int *)A(64);
  int32_t *wl = (int32_t *)A(nev * 12 + 16);
  bool ok = evk && cnt && nruns && nev_dev && eoff && ncl && nmem && cl_ev && cl_type && cl_pc && cl_sh && cl_d &&
            cl_moff && mem && clo && mo && qk && qk2 && qi && qo && qctr && wl;
  for (auto q : bd) ok = ok && *q;
  for (auto q : bi) ok = ok && *q;
  for (auto q : bu) ok = ok && *q;
  if (!ok) return set_error(c, CRV_E_HIP, "graph: finder buffer allocation failed" + ar.note());
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, ev, evk, cnt, nruns, (int)n, s));
  void *tmp = A(tb);
  if (!tmp) return set_error(c, CRV_E_HIP, "graph: run-length buffer allocation failed");
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp, tb, ev, evk, cnt, nruns, (int)n, s));
  k_fix_events<<<nblocks(nev + 1, 256), 256, 0, s>>>(nruns, pl.sizes, n, cnt, nev_dev);
  CRV_HIP(c, hipGetLastError());
  int rc = widen_scan_fixed(c, ar, cnt, eoff, nev, s);
  if (rc) return rc;
  Ev e{bar, sipm, seq};
  const ClSlots slots{cl_ev, cl_type, cl_pc, cl_sh, cl_d};
  const int grid = (int)std::min<int64_t>(nev, 16384);
  k_zero_i32<<<1, 64, 0, s>>>(qctr, 16);
  k_zero_i32<<<nblocks(nev + 1, 256), 256, 0, s>>>((int32_t *)qk, nev + 1);
  k_zero_i32<<<nblocks(nev + 1, 256), 256, 0, s>>>(ncl, nev + 1);   // events past the call's find nothing
  k_zero_i32<<<nblocks(nev + 1, 256), 256, 0, s>>>(nmem, nev + 1);
  k_block_keys<<<nblocks(nev, 256), 256, 0, s>>>(nev, eoff, (int64_t)257, qk, qi, qctr + 1, wl, nev_dev);
  CRV_HIP(c, hipGetLastError());
  size_t qtb = 0;
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, qtb, qk, qk2, qi, qo, (int)nev, 0, 16, c->aux[1]));
  void *qtmp = A(qtb);
  if (!qtmp) return set_error(c, CRV_E_HIP, "graph: queue sort buffer allocation failed");
  CRV_HIP(c, hipEventRecord(c->fork[0], s));
  CRV_HIP(c, hipStreamWaitEvent(c->aux[0], c->fork[0], 0));
  CRV_HIP(c, hipStreamWaitEvent(c->aux[1], c->fork[0], 0));
  k_coinc_wave<64><<<grid, 64, 0, s>>>(c->p, c->t, wl, qctr + 2, eoff, evk, time, pe, e, ncl, nmem, slots, cl_moff,
                                       mem, c->flags);
  const hipError_t e0 = hipGetLastError();
  k_coinc_wave<128><<<grid, 64, 0, c->aux[0]>>>(c->p, c->t, wl + nev, qctr + 3, eoff, evk, time, pe, e, ncl, nmem,
                                                slots, cl_moff, mem, c->flags);
  const hipError_t e1 = hipGetLastError();
  k_coinc_wave<256><<<std::min(grid, 4096), 64, 0, c->aux[1]>>>(c->p, c->t, wl + 2 * nev, qctr + 4, eoff, evk, time,
                                                                pe, e, ncl, nmem, slots, cl_moff, mem, c->flags);
  const hipError_t e2 = hipGetLastError();
  hipError_t e3 = hipcub::DeviceRadixSort::SortPairsDescending(qtmp, qtb, qk, qk2, qi, qo, (int)nev, 0, 16, c->aux[1]);
  if (e3 == hipSuccess) {
    k_coinc_block<<<2048, BT, 0, c->aux[1]>>>(c->p, c->t, nev, eoff, evk, time, pe, e, b, ncl, nmem, slots, cl_moff,
                                             mem, (int64_t)257, qo, qk2, qctr, c->flags);
    e3 = hipGetLastError();
  }
  CRV_HIP(c, hipEventRecord(c->fork[1], c->aux[0]));
  CRV_HIP(c, hipEventRecord(c->fork[2], c->aux[1]));
  CRV_HIP(c, hipStreamWaitEvent(s, c->fork[1], 0));
  CRV_HIP(c, hipStreamWaitEvent(s, c->fork[2], 0));
  CRV_HIP(c, e0);
  CRV_HIP(c, e1);
  CRV_HIP(c, e2);
  CRV_HIP(c, e3);
  if ((rc = widen_scan_fixed(c, ar, ncl, clo, nev, s))) return rc;
  if ((rc = widen_scan_fixed(c, ar, nmem, mo, nev, s))) return rc;
  k_cluster_sizes<<<1, 1, 0, s>>>(clo, mo, nev, pl.skip, pl.sizes);
  crv_clusters C = pl.C;
  C.capacity = n;
  C.mem_capacity = n;
  k_coinc_out<<<nblocks(nev, 256), 256, 0, s>>>(nev, eoff, clo, mo, ncl, cl_ev, cl_type, cl_pc, cl_sh, cl_d, cl_moff,
                                                mem, C, nev_dev);
  CRV_HIP(c, hipGetLastError());
  return CRV_OK;
}

}  // namespace crv
