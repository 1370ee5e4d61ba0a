// Coincidence finder: pulses -> coincidence clusters (FP64): one wavefront per event of up to 256 pulses (the
// event's hits in LDS), one lane per larger event.
//
// Follows the reference's reco/coinc.py:112-392 and reco/coinc_cpu.py:11-51 step by step: hits from first pulses
// with the fibre-length time correction and partner-status PE compensation; filterHits per readout side (same and
// adjacent counter within max_dt_adjacent); greedy findClusters (repeated passes in list order, a hit joins through
// any member within max(maxd) in x and max_dt in time); checkCoincidence over layer combinations (time window,
// slopes, slope differences, maxDistance widened to the x gaps of passing combinations, kept hits in pulse index
// order); findClusters again across sides; opposite-side attachment (nearest cluster per candidate, one pulse per
// SiPM in (cluster, bar, sipm, dt, candidate) order); cluster moments as sequential sums in member order.
// Lists live in per-event scratch ranges of the pulse arrays.
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

struct Scr {  // per-pulse scratch (indexed from the event's first pulse)
  double *t, *pe, *maxd, *wmaxd, *dt;
  int32_t *hl, *filt, *rem, *items, *cl_of, *coinc, *bylayer, *chosen_c, *chosen_k, *cl2_off, *cl2_items;
  uint8_t *kept;
};

struct Ev {  // what the finder needs per pulse: bar, sipm, seq, idx and the derived hit fields
  const int32_t *bar, *sipm;
  const int64_t *seq;
};

// link test of reco/coinc.py:142-146
__device__ __forceinline__ bool linked(const Tables &t, const Ev &e, const Scr &s, int h, int c, const double *maxd,
                                       double max_dt, bool opposite, double dt_opp) {
  double md = fmax(maxd[h], maxd[c]);
  int sh = e.sipm[h] % 2, sc = e.sipm[c] % 2;
  double mt = max_dt + ((opposite && sh != sc) ? dt_opp : 0.0);
  return fabs(t.bx[e.bar[h]] - t.bx[e.bar[c]]) <= md && fabs(s.t[h] - s.t[c]) < mt;
}

// greedy findClusters of list[0..n): writes members in join order to out[0..n) and cluster starts to off;
// returns the number of clusters. rem is scratch of size n.
__device__ int find_clusters(const Tables &t, const Ev &e, const Scr &s, const int32_t *list, int n,
                             const double *maxd, double max_dt, bool opposite, double dt_opp, int32_t *rem,
                             int32_t *out, int32_t *off) {
  for (int i = 0; i < n; ++i) rem[i] = list[i];
  int nrem = n, nout = 0, ncl = 0;
  while (nrem > 0) {
    int cs = nout;
    off[ncl] = cs;
    int last = -1;
    while (last != nout - cs) {
      last = nout - cs;
      int w = 0;
      for (int i = 0; i < nrem; ++i) {
        int h = rem[i];
        bool join = nout == cs;
        for (int m = cs; m < nout && !join; ++m) join = linked(t, e, s, h, out[m], maxd, max_dt, opposite, dt_opp);
        if (join) out[nout++] = h; else rem[w++] = h;
      }
      nrem = w;
    }
    ++ncl;
  }
  off[ncl] = nout;
  return ncl;
}

// checkCoincidence (reco/coinc.py:184-207) of cluster members m[0..n): appends kept hits (pulse index order) to
// coinc with their widened max distance; returns the new coinc length.
__device__ int check_coincidence(const crv_params &p, const Tables &t, const Ev &e, const Scr &s, const int32_t *m,
                                 int n, int32_t *coinc, int nco, double *cmaxd, int32_t *byl, double *wmaxd,
                                 uint8_t *kept) {
  if (n == 0) return nco;
  int lay_off[5] = {0, 0, 0, 0, 0};
  int nmin = 1 << 30, nmax = -1;
  for (int i = 0; i < n; ++i) {
    Cut c = cut_of(t, t.sector[e.bar[m[i]]]);
    int cl = (int)c.coincidence_layers;
    nmin = cl < nmin ? cl : nmin;
    nmax = cl > nmax ? cl : nmax;
  }
  // by_layer lists in cluster order; wmaxd / kept indexed by position in byl
  int pos = 0;
  for (int l = 0; l < 4; ++l) {
    lay_off[l] = pos;
    for (int i = 0; i < n; ++i)
      if (t.layer[e.bar[m[i]]] == l) {
        byl[pos] = m[i];
        wmaxd[pos] = cut_of(t, t.sector[e.bar[m[i]]]).final_max_distance;
        kept[pos] = 0;
        ++pos;
      }
  }
  lay_off[4] = pos;
  int nonempty = 0;
  for (int l = 0; l < 4; ++l) nonempty += lay_off[l + 1] > lay_off[l];
  if (n > p.big_cluster_threshold && nonempty >= nmin) {
    for (int i = 0; i < n; ++i) {
      coinc[nco] = m[i];
      cmaxd[nco] = s.maxd[m[i]];
      ++nco;
    }
    return nco;
  }
  for (int nl = 2; nl <= 4; ++nl) {
    bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    // layer combinations in lexicographic order
    int L[4];
    for (int i = 0; i < nl; ++i) L[i] = i;
    for (;;) {
      bool empty = false;
      for (int i = 0; i < nl; ++i) empty |= lay_off[L[i] + 1] == lay_off[L[i]];
      if (!empty) {
        int idx[4];
        for (int i = 0; i < nl; ++i) idx[i] = lay_off[L[i]];
        for (;;) {
          // combo = byl[idx[0..nl)]
          bool skip = true;
          double mdt = -INFINITY, lo = INFINITY, hi = -INFINITY, msd = -INFINITY;
          double tmin = INFINITY, tmax = -INFINITY;
          for (int i = 0; i < nl; ++i) {
            int h = byl[idx[i]];
            Cut c = cut_of(t, t.sector[e.bar[h]]);
            if (!((int)c.coincidence_layers > nl)) skip = false;
            mdt = fmax(mdt, c.max_dt);
            lo = fmin(lo, c.min_slope);
            hi = fmax(hi, c.max_slope);
            msd = fmax(msd, c.max_slope_diff);
            tmin = fmin(tmin, s.t[h]);
            tmax = fmax(tmax, s.t[h]);
          }
          bool pass = !skip && !(tmax - tmin > mdt);
          double sl[3];
          for (int i = 0; pass && i + 1 < nl; ++i) {
            int a = byl[idx[i]], b = byl[idx[i + 1]];
            double v = (t.bx[e.bar[b]] - t.bx[e.bar[a]]) / (t.by[e.bar[b]] - t.by[e.bar[a]]);
            if (v < lo || v > hi) pass = false;
            sl[i] = v;
          }
          if (pass && nl > 2 && fabs(sl[0] - sl[1]) > msd) pass = false;
          if (pass && nl > 3 && (fabs(sl[0] - sl[2]) > msd || fabs(sl[1] - sl[2]) > msd)) pass = false;
          if (pass) {
            // stable sort of the combo by x, then widen to the neighbour gaps
            int ord[4];
            for (int i = 0; i < nl; ++i) ord[i] = i;
            for (int i = 1; i < nl; ++i) {
              int v = ord[i], j = i - 1;
              double xv = t.bx[e.bar[byl[idx[v]]]];
              while (j >= 0 && t.bx[e.bar[byl[idx[ord[j]]]]] > xv) {
                ord[j + 1] = ord[j];
                --j;
              }
              ord[j + 1] = v;
            }
            for (int i = 0; i < nl; ++i) {
              int q = idx[ord[i]];
              double x = t.bx[e.bar[byl[q]]];
              if (i > 0) wmaxd[q] = fmax(wmaxd[q], x - t.bx[e.bar[byl[idx[ord[i - 1]]]]]);
              if (i + 1 < nl) wmaxd[q] = fmax(wmaxd[q], t.bx[e.bar[byl[idx[ord[i + 1]]]]] - x);
              kept[q] = 1;
            }
          }
          // next product element (last list fastest)
          int k = nl - 1;
          while (k >= 0) {
            if (++idx[k] < lay_off[L[k] + 1]) break;
            idx[k] = lay_off[L[k]];
            --k;
          }
          if (k < 0) break;
        }
      }
      // next layer combination
      int k = nl - 1;
      while (k >= 0 && L[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++L[k];
      for (int i = k + 1; i < nl; ++i) L[i] = L[i - 1] + 1;
    }
  }
  // kept hits in pulse index order (the event's pulse rows are in index order)
  int first = nco;
  for (int q = 0; q < pos; ++q) {
    if (!kept[q]) continue;
    int h = byl[q];
    int at = nco;
    while (at > first && coinc[at - 1] > h) {
      coinc[at] = coinc[at - 1];
      cmaxd[at] = cmaxd[at - 1];
      --at;
    }
    coinc[at] = h;
    cmaxd[at] = wmaxd[q];
    ++nco;
  }
  return nco;
}


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
  double tot = 0.0, mincp = INFINITY;
  for (int i = 0; i < mn; ++i) {
    int h = mem[i];
    tot = tot + A.pe(h);
    mincp = fmin(mincp, cut_of(t, t.sector[A.bar(h)]).min_cluster_pes);
  }
  if (tot < mincp) return false;
  double pos[3];
  for (int ax = 0; ax < 3; ++ax) {
    double acc = 0.0;
    for (int i = 0; i < mn; ++i) {
      int h = mem[i];
      acc = acc + t.bpos[A.bar(h) * 3 + ax] * A.pe(h);
    }
    pos[ax] = acc / tot;
  }
  double sxy = 0.0, sx = 0.0, sy = 0.0, syy = 0.0;
  double st[2] = {0.0, 0.0}, spe[2] = {0.0, 0.0}, tmn[2] = {0.0, 0.0}, tmx[2] = {0.0, 0.0};
  int nhs[2] = {0, 0};
  double cmin = INFINITY, cmax = -INFINITY;
  for (int i = 0; i < mn; ++i) {
    int h = mem[i];
    double pe = A.pe(h), x = t.bx[A.bar(h)], y = t.by[A.bar(h)], tt = A.tm(h);
    sxy = sxy + pe * x * y;
    sx = sx + pe * x;
    sy = sy + pe * y;
    syy = syy + pe * y * y;
    cmin = fmin(cmin, tt);
    cmax = fmax(cmax, tt);
  }
  for (int sd = 0; sd < 2; ++sd) {
    double ps = 0.0, pt = 0.0;
    for (int i = 0; i < mn; ++i) {
      int h = mem[i];
      if (A.sipm(h) % 2 != sd || !(A.pe(h) > 0)) continue;
      double tt = A.tm(h);
      if (nhs[sd] == 0) {
        tmn[sd] = tt;
        tmx[sd] = tt;
      } else {
        tmn[sd] = fmin(tmn[sd], tt);
        tmx[sd] = fmax(tmx[sd], tt);
      }
      ++nhs[sd];
      ps = ps + A.pe(h);
      pt = pt + tt * A.pe(h);
    }
    spe[sd] = ps;
    st[sd] = ps > 0 ? pt / ps : 0.0;
  }
  double slope = (tot * sxy - sx * sy) / (tot * syy - sy * sy);
  int ld = t.ldir[t.sector[A.bar(mem[0])]];
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

struct SerialAcc {
  const Ev &E;
  const Scr &s;
  __device__ double pe(int h) const { return s.pe[h]; }
  __device__ double tm(int h) const { return s.t[h]; }
  __device__ int bar(int h) const { return E.bar[h]; }
  __device__ int sipm(int h) const { return E.sipm[h]; }
};

__global__ void k_coinc(crv_params p, Tables t, int64_t nev, const int64_t *__restrict__ eoff,
                        const int64_t *__restrict__ evk, const double *__restrict__ ptime,
                        const double *__restrict__ ppe, Ev e, Scr s0, int32_t *__restrict__ ncl,
                        int32_t *__restrict__ nmem, int64_t *__restrict__ cl_ev, int32_t *__restrict__ cl_type,
                        int32_t *__restrict__ cl_poscalc, int32_t *__restrict__ cl_sidehits,
                        double *__restrict__ cl_d /* [slot * 12] */, int32_t *__restrict__ cl_moff,
                        int32_t *__restrict__ mem, int64_t skip_n, int *__restrict__ flags) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nev) return;
  const int64_t a = eoff[g], n = eoff[g + 1] - a;
  if (n <= skip_n) return;  // taken by the wavefront kernel
  const int64_t evkey = evk[g];
  // event-local views: scratch ranges start at a + g, so every event owns n + 1 entries
  const int64_t sb = a + g;
  Ev E{e.bar + a, e.sipm + a, e.seq + a};
  Scr s = s0;
  s.t += sb; s.pe += sb; s.maxd += sb; s.wmaxd += sb; s.dt += sb;
  s.hl += sb; s.filt += sb; s.rem += sb; s.items += sb; s.cl_of += sb; s.coinc += sb; s.bylayer += sb;
  s.chosen_c += sb; s.chosen_k += sb; s.cl2_off += sb; s.cl2_items += sb; s.kept += sb;
  int32_t *cl_moff_e = cl_moff + sb;  // cluster member offsets
  const ClSlots slots{cl_ev, cl_type, cl_poscalc, cl_sidehits, cl_d};
  int32_t *mem_e = mem + a;
  int nclus = 0, nm = 0;
  // hits: per sector type, in pulse order (types ascending)
  for (int ti = 0; ti < t.ntypes; ++ti) {
    const int type = t.type_id[ti];
    const double P0 = t.tp0[ti], P1 = t.tp1[ti], half = t.thalf[ti], tmaxdt = t.tmaxdt[ti];
    const bool both = t.tside0[ti] && t.tside1[ti];
    int nh = 0;
    for (int k = 0; k < n; ++k) {
      if (p.ignore_secondary_pulses && E.seq[k] > 0) continue;
      int bar = E.bar[k];
      if (bar < 0 || bar >= t.nbars || E.sipm[k] < 0 || E.sipm[k] > 3 || bar * 4 + 3 >= t.n_channels) {
        atomicAdd(&flags[F_BADCHAN], 1);
        continue;
      }
      int sec = t.sector[bar];
      if (!t.hascuts[sec] || t.stype[sec] != type) continue;
      int sipm = E.sipm[k];
      double tt = ptime[a + k];
      if (sipm % 2 == 0) tt += (t.ps0[sec] - P0) / p.fiber_signal_speed;
      else tt += (P1 - t.ps1[sec]) / p.fiber_signal_speed;
      double pe = ppe[a + k];
      int other = bar * 4 + sipm + (sipm < 2 ? 2 : -2);
      if (t.status[other] & p.compensate_mask) pe *= 2.0;
      s.t[k] = tt;
      s.pe[k] = pe;
      s.maxd[k] = cut_of(t, sec).initial_max_distance;
      s.hl[nh++] = k;
    }
    if (nh == 0) continue;
    int nco = 0, nfilt = 0;
    for (int side = 0; side < 2; ++side) {
      int f0 = nfilt;
      for (int ii = 0; ii < nh; ++ii) {
        int h = s.hl[ii];
        if (E.sipm[h] % 2 != side) continue;
        Cut ch = cut_of(t, t.sector[E.bar[h]]);
        int lh = t.layer[E.bar[h]], cnh = t.cit[E.bar[h]];
        double same = 0.0, adj1 = 0.0, adj2 = 0.0;
        for (int jj = 0; jj < nh; ++jj) {
          int q = s.hl[jj];
          if (E.sipm[q] % 2 != side || t.layer[E.bar[q]] != lh || fabs(s.t[q] - s.t[h]) > ch.max_dt_adjacent) continue;
          int d = t.cit[E.bar[q]] - cnh;
          same += d == 0 ? s.pe[q] : 0.0;
          adj1 += d == -1 ? s.pe[q] : 0.0;
          adj2 += d == 1 ? s.pe[q] : 0.0;
        }
        if (same + adj1 >= ch.pe_threshold || same + adj2 >= ch.pe_threshold) s.filt[nfilt++] = h;
      }
      // clusters of this side's filtered hits, then the coincidence check per cluster
      int ncs = find_clusters(t, E, s, s.filt + f0, nfilt - f0, s.maxd, tmaxdt, false, 0.0, s.rem, s.items,
                              s.cl_of);
      for (int c = 0; c < ncs; ++c) {
        int c0 = s.cl_of[c], c1 = s.cl_of[c + 1];
        nco = check_coincidence(p, t, E, s, s.items + c0, c1 - c0, s.coinc, nco, s.dt, s.bylayer, s.wmaxd, s.kept);
      }
    }
    const double dt_opp = 2.0 * half / p.fiber_signal_speed;
    // coinc hits carry their widened max distance in s.dt; cluster across sides
    for (int i = 0; i < nco; ++i) s.wmaxd[s.coinc[i]] = s.dt[i];
    int ncl2 = find_clusters(t, E, s, s.coinc, nco, s.wmaxd, tmaxdt, both, dt_opp, s.rem, s.cl2_items, s.cl2_off);
    int nitems = s.cl2_off[ncl2];
    // attachment (reco/coinc_cpu.py:11-51): appended per cluster after its members
    int nchosen = 0;
    if (both && p.add_opposite_side_pulses && nitems > 0 && nfilt > 0) {
      const double mdt = tmaxdt + dt_opp;
      for (int ci = 0; ci < nfilt; ++ci) {
        int h = s.filt[ci];
        if (E.seq[h] != 0) continue;
        bool member = false;
        for (int m = 0; m < nitems && !member; ++m) member = s.cl2_items[m] == h;
        if (member) continue;
        int hb = E.bar[h], hs = E.sipm[h] % 2;
        int best_c = -1;
        double best_dt = 0.0;
        for (int c = 0; c < ncl2; ++c)
          for (int m = s.cl2_off[c]; m < s.cl2_off[c + 1]; ++m) {
            int an = s.cl2_items[m];
            if (E.seq[an] != 0 || E.bar[an] != hb || hs != 1 - E.sipm[an] % 2) continue;
            double dt = fabs(s.t[h] - s.t[an]);
            if (!(dt < mdt)) continue;
            bool occ = false;
            for (int q = s.cl2_off[c]; q < s.cl2_off[c + 1] && !occ; ++q)
              occ = E.bar[s.cl2_items[q]] == hb && E.sipm[s.cl2_items[q]] % 2 == hs;
            if (occ) continue;
            if (best_c < 0 || dt < best_dt || (dt == best_dt && c < best_c)) {
              best_c = c;
              best_dt = dt;
            }
          }
        if (best_c < 0) continue;
        // insert in (cluster, bar, sipm, dt, candidate) order
        int at = nchosen;
        while (at > 0) {
          int pc = s.chosen_k[at - 1], ph = s.filt[s.chosen_c[at - 1]];
          double pdt = s.dt[at - 1];
          bool greater = pc > best_c ||
                         (pc == best_c && (E.bar[ph] > hb || (E.bar[ph] == hb && (E.sipm[ph] > E.sipm[h] ||
                         (E.sipm[ph] == E.sipm[h] && (pdt > best_dt || (pdt == best_dt && s.chosen_c[at - 1] > ci)))))));
          if (!greater) break;
          s.chosen_k[at] = s.chosen_k[at - 1];
          s.chosen_c[at] = s.chosen_c[at - 1];
          s.dt[at] = s.dt[at - 1];
          --at;
        }
        s.chosen_k[at] = best_c;
        s.chosen_c[at] = ci;
        s.dt[at] = best_dt;
        ++nchosen;
      }
    }
    // clusters with their attachments, moments
    int q = 0;
    for (int c = 0; c < ncl2; ++c) {
      int base = nm;
      for (int m = s.cl2_off[c]; m < s.cl2_off[c + 1]; ++m) mem_e[nm++] = s.cl2_items[m];
      int lastb = -1, lasts = -1;
      for (; q < nchosen && s.chosen_k[q] == c; ++q) {
        int h = s.filt[s.chosen_c[q]];
        if (E.bar[h] == lastb && E.sipm[h] == lasts) continue;  // one pulse per (cluster, bar, sipm)
        lastb = E.bar[h];
        lasts = E.sipm[h];
        mem_e[nm++] = h;
      }
      const int mn = nm - base;
      if (!cluster_emit(p, t, SerialAcc{E, s}, mem_e + base, mn, type, P0, P1, evkey, a + nclus, slots)) {
        nm = base;
        continue;
      }
      cl_moff_e[nclus] = base;
      ++nclus;
    }
  }
  cl_moff_e[nclus] = nm;
  ncl[g] = nclus;
  nmem[g] = nm;
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
};

template <int WMAX>
struct WaveAcc {
  const WS<WMAX> &S;
  __device__ double pe(int h) const { return S.pe[h]; }
  __device__ double tm(int h) const { return S.t[h]; }
  __device__ int bar(int h) const { return S.bar[h]; }
  __device__ int sipm(int h) const { return S.sipm[h]; }
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

__device__ __forceinline__ void lds_fmax(double *addr, double v) {
  unsigned long long *a = (unsigned long long *)addr;
  unsigned long long old = *a, assumed;
  do {
    assumed = old;
    if (!(v > __longlong_as_double((long long)assumed))) return;
    old = atomicCAS(a, assumed, (unsigned long long)__double_as_longlong(v));
  } while (old != assumed);
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

// checkCoincidence of the cluster mem[0..n) (pulses): kept hits appended to coinc in pulse order with their widened
// max distance; returns the new coinc length
template <int WMAX>
__device__ int wave_check(const crv_params &p, const Tables &t, WS<WMAX> &S, const uint8_t *mem, int n, int nco) {
  constexpr int WNW = WMAX / 64;
  const int lane = threadIdx.x;
  if (n == 0) return nco;
  int nmin = 1 << 30, nmax = -1;
  for (int i = lane; i < n; i += 64) {
    const int cl = (int)cut_of(t, S.sec[mem[i]]).coincidence_layers;
    nmin = min(nmin, cl);
    nmax = max(nmax, cl);
  }
  nmin = wave_min(nmin);
  nmax = wave_max(nmax);
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
        S.wq[q] = cut_of(t, S.sec[mem[i]]).final_max_distance;
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
  for (int nl = 2; nl <= 4; ++nl) {
    const bool need = nl == 2 ? nmin == 2 : nl == 3 ? (nmin <= 3 && 3 <= nmax) : nmax == 4;
    if (!need) continue;
    int L[4];
    for (int i = 0; i < nl; ++i) L[i] = i;
    for (;;) {
      int64_t P = 1;
      for (int i = 0; i < nl; ++i) P *= lay_off[L[i] + 1] - lay_off[L[i]];
      for (int64_t cidx = lane; cidx < P; cidx += 64) {
        int idx[4];
        if (P <= INT32_MAX) {  // 32-bit decode (64-bit division is a long software sequence)
          uint32_t rest = (uint32_t)cidx;
          for (int i = nl - 1; i >= 0; --i) {
            const uint32_t sz = (uint32_t)(lay_off[L[i] + 1] - lay_off[L[i]]);
            idx[i] = lay_off[L[i]] + (int)(rest % sz);
            rest /= sz;
          }
        } else {
          int64_t rest = cidx;
          for (int i = nl - 1; i >= 0; --i) {
            const int sz = lay_off[L[i] + 1] - lay_off[L[i]];
            idx[i] = lay_off[L[i]] + (int)(rest % sz);
            rest /= sz;
          }
        }
        bool skip = true;
        double mdt = -INFINITY, lo = INFINITY, hi = -INFINITY, msd = -INFINITY;
        double tmin = INFINITY, tmax = -INFINITY;
        for (int i = 0; i < nl; ++i) {
          const int h = S.byl[idx[i]];
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
            lds_fmax(&S.wq[q], w);
            S.kept[q] = 1;
          }
        }
      }
      int k = nl - 1;
      while (k >= 0 && L[k] == 4 - nl + k) --k;
      if (k < 0) break;
      ++L[k];
      for (int i = k + 1; i < nl; ++i) L[i] = L[i - 1] + 1;
    }
  }
  __syncthreads();
  for (int q = lane; q < pos; q += 64)
    if (S.kept[q]) {
      const int h = S.byl[q];
      S.wm[h] = S.wq[q];  // a pulse is in one cluster: the slot by pulse is free
      atomicOr((unsigned long long *)&S.kmask[h >> 6], 1ull << (h & 63));
    }
  __syncthreads();
  // kept hits in pulse order
  for (int w = 0; w < WNW; ++w) {
    uint64_t r = S.kmask[w];
    while (r) {
      const int b = __ffsll((unsigned long long)r) - 1;
      r &= r - 1;
      const int h = w * 64 + b;
      if (lane == 0) {
        S.coinc[nco] = (uint8_t)h;
        S.cmaxd[nco] = S.wm[h];
      }
      ++nco;
    }
  }
  __syncthreads();
  return nco;
}

template <int WMAX>
__global__ __launch_bounds__(64) void k_coinc_wave(crv_params p, Tables t, int64_t nev, const int64_t *__restrict__ eoff,
                                                   const int64_t *__restrict__ evk, const double *__restrict__ ptime,
                                                   const double *__restrict__ ppe, Ev e, int32_t *__restrict__ ncl,
                                                   int32_t *__restrict__ nmem, ClSlots slots,
                                                   int32_t *__restrict__ cl_moff, int32_t *__restrict__ mem,
                                                   int64_t min_n, int *__restrict__ flags) {
  constexpr int WNW = WMAX / 64;
  __shared__ WS<WMAX> S;
  const int lane = threadIdx.x;
  for (int64_t g = blockIdx.x; g < nev; g += gridDim.x) {
    const int64_t a = eoff[g], n64 = eoff[g + 1] - a;
    if (n64 > WMAX || n64 < min_n) continue;  // another size class takes it
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
            const int sec = t.sector[bar];
            if (t.hascuts[sec] && t.stype[sec] == type) {
              double tt = ptime[a + k];
              if (sp % 2 == 0) tt += (t.ps0[sec] - P0) / p.fiber_signal_speed;
              else tt += (P1 - t.ps1[sec]) / p.fiber_signal_speed;
              double pe = ppe[a + k];
              const int other = bar * 4 + sp + (sp < 2 ? 2 : -2);
              if (t.status[other] & p.compensate_mask) pe *= 2.0;
              S.t[k] = tt;
              S.pe[k] = pe;
              S.maxd[k] = cut_of(t, sec).initial_max_distance;
              S.x[k] = t.bx[bar];
              S.y[k] = t.by[bar];
              S.sec[k] = (int16_t)sec;
              S.layer[k] = (uint8_t)t.layer[bar];
              S.cit[k] = (int16_t)t.cit[bar];
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
        // order (cluster, bar, sipm, dt, candidate): rank of each chosen candidate among the chosen
        for (int base = 0; base < nfilt; base += 64) {
          const int ci = base + lane;
          const bool chosen = ci < nfilt && S.chk[ci] >= 0;
          if (chosen) {
            const int k = S.chk[ci], h = S.filt[ci];
            const int b = S.bar[h], sp = S.sipm[h];
            const double dt = S.cdt[ci];
            int rank = 0;
            for (int cj = 0; cj < nfilt; ++cj) {
              const int kj = S.chk[cj];
              if (kj < 0 || cj == ci) continue;
              const int hj = S.filt[cj];
              const int bj = S.bar[hj], sj = S.sipm[hj];
              const double dj = S.cdt[cj];
              const bool less = kj < k || (kj == k && (bj < b || (bj == b && (sj < sp || (sj == sp && (dj < dt ||
                                (dj == dt && cj < ci)))))));
              rank += less;
            }
            S.sk[rank] = (int16_t)k;
            S.sc[rank] = (uint8_t)ci;
          }
          nchosen += __popcll(__ballot(chosen));
        }
        __syncthreads();
      }
      // members of each second-level cluster (items, then its attachments, one per (bar, sipm)); min-PE test
      for (int base = 0; base < ncl2; base += 64) {
        const int c = base + lane;
        if (c < ncl2) {
          int mc = S.off2[c + 1] - S.off2[c];
          double tot = 0.0, mincp = INFINITY;
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) {
            const int h = S.cl2[m];
            tot = tot + S.pe[h];
            mincp = fmin(mincp, cut_of(t, S.sec[h]).min_cluster_pes);
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
            mincp = fmin(mincp, cut_of(t, S.sec[h]).min_cluster_pes);
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
          int32_t *mm = mem_e + mbase;
          int w = 0;
          for (int m = S.off2[c]; m < S.off2[c + 1]; ++m) mm[w++] = S.cl2[m];
          int lastb = -1, lasts = -1;
          for (int q = 0; q < nchosen; ++q) {
            if (S.sk[q] != c) continue;
            const int h = S.filt[S.sc[q]];
            if (S.bar[h] == lastb && S.sipm[h] == lasts) continue;
            lastb = S.bar[h];
            lasts = S.sipm[h];
            mm[w++] = h;
          }
          cluster_emit(p, t, WaveAcc<WMAX>{S}, mm, mc, type, P0, P1, evkey, a + cslot, slots);
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

__global__ void k_coinc_out(int64_t nev, const int64_t *__restrict__ eoff, const int64_t *__restrict__ clo,
                            const int64_t *__restrict__ mo, const int32_t *__restrict__ ncl,
                            const int64_t *__restrict__ cl_ev, const int32_t *__restrict__ cl_type,
                            const int32_t *__restrict__ cl_poscalc, const int32_t *__restrict__ cl_sidehits,
                            const double *__restrict__ cl_d, const int32_t *__restrict__ cl_moff,
                            const int32_t *__restrict__ mem, crv_clusters o) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
  Scr sc;
  double **dp[] = {&sc.t, &sc.pe, &sc.maxd, &sc.wmaxd, &sc.dt};
  const int64_t ns = n + nev + 2;
  for (int i = 0; i < 5; ++i) *dp[i] = (double *)ws(c, 96 + i, ns * 8);
  int32_t **ip[] = {&sc.hl, &sc.filt, &sc.rem, &sc.items, &sc.cl_of, &sc.coinc, &sc.bylayer, &sc.chosen_c,
                    &sc.chosen_k, &sc.cl2_off, &sc.cl2_items};
  for (int i = 0; i < 11; ++i) *ip[i] = (int32_t *)ws(c, 101 + i, ns * 4);
  sc.kept = (uint8_t *)ws(c, 112, ns + 8);
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
  if (!sc.kept || !ncl || !nmem || !cl_ev || !cl_type || !cl_pc || !cl_sh || !cl_d || !cl_moff || !mem || !clo || !mo)
    return ws_fail(c);
  Ev e{bar, sipm, seq};
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  kmark(c, 3, 0, s);
  {  // size classes: wavefront kernels for events of <= 128 and <= 256 pulses, one lane per larger event
    const ClSlots slots{cl_ev, cl_type, cl_pc, cl_sh, cl_d};
    const int grid = (int)std::min<int64_t>(nev, 16384);
    k_coinc_wave<128><<<grid, 64, 0, s>>>(c->p, c->t, nev, eoff, evk, time, pe, e, ncl, nmem, slots, cl_moff, mem,
                                          (int64_t)0, c->flags);
    CRV_HIP(c, hipGetLastError());
    k_coinc_wave<256><<<std::min(grid, 4096), 64, 0, s>>>(c->p, c->t, nev, eoff, evk, time, pe, e, ncl, nmem, slots,
                                                          cl_moff, mem, (int64_t)129, c->flags);
    CRV_HIP(c, hipGetLastError());
  }
  k_coinc<<<nblocks(nev, 64), 64, 0, s>>>(c->p, c->t, nev, eoff, evk, time, pe, e, sc, ncl, nmem, cl_ev, cl_type,
                                          cl_pc, cl_sh, cl_d, cl_moff, mem, (int64_t)256, c->flags);
  CRV_HIP(c, hipGetLastError());
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

}  // namespace crv
