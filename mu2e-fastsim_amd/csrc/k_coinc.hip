// Coincidence finder: pulses -> coincidence clusters (FP64), one lane per event.
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

struct ClOut {
  crv_clusters o;
};

__global__ void k_coinc(crv_params p, Tables t, int64_t nev, const int64_t *__restrict__ eoff,
                        const int64_t *__restrict__ evk, const double *__restrict__ ptime,
                        const double *__restrict__ ppe, Ev e, Scr s0, int32_t *__restrict__ ncl,
                        int32_t *__restrict__ nmem, int64_t *__restrict__ cl_ev, int32_t *__restrict__ cl_type,
                        int32_t *__restrict__ cl_poscalc, int32_t *__restrict__ cl_sidehits,
                        double *__restrict__ cl_d /* [slot * 12] */, int32_t *__restrict__ cl_moff,
                        int32_t *__restrict__ mem, int *__restrict__ flags) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nev) return;
  const int64_t a = eoff[g], n = eoff[g + 1] - a;
  const int64_t evkey = evk[g];
  // event-local views: scratch ranges start at a + g, so every event owns n + 1 entries
  const int64_t sb = a + g;
  Ev E{e.bar + a, e.sipm + a, e.seq + a};
  Scr s = s0;
  s.t += sb; s.pe += sb; s.maxd += sb; s.wmaxd += sb; s.dt += sb;
  s.hl += sb; s.filt += sb; s.rem += sb; s.items += sb; s.cl_of += sb; s.coinc += sb; s.bylayer += sb;
  s.chosen_c += sb; s.chosen_k += sb; s.cl2_off += sb; s.cl2_items += sb; s.kept += sb;
  int32_t *cl_moff_e = cl_moff + sb;  // cluster member offsets
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
      double tot = 0.0, mincp = INFINITY;
      for (int i = 0; i < mn; ++i) {
        int h = mem_e[base + i];
        tot = tot + s.pe[h];
        mincp = fmin(mincp, cut_of(t, t.sector[E.bar[h]]).min_cluster_pes);
      }
      if (tot < mincp) {
        nm = base;
        continue;
      }
      double pos[3];
      for (int ax = 0; ax < 3; ++ax) {
        double acc = 0.0;
        for (int i = 0; i < mn; ++i) {
          int h = mem_e[base + i];
          acc = acc + t.bpos[E.bar[h] * 3 + ax] * s.pe[h];
        }
        pos[ax] = acc / tot;
      }
      double sxy = 0.0, sx = 0.0, sy = 0.0, syy = 0.0;
      double st[2] = {0.0, 0.0}, spe[2] = {0.0, 0.0}, tmn[2] = {0.0, 0.0}, tmx[2] = {0.0, 0.0};
      int nhs[2] = {0, 0};
      double cmin = INFINITY, cmax = -INFINITY;
      for (int i = 0; i < mn; ++i) {
        int h = mem_e[base + i];
        double pe = s.pe[h], x = t.bx[E.bar[h]], y = t.by[E.bar[h]], tt = s.t[h];
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
          int h = mem_e[base + i];
          if (E.sipm[h] % 2 != sd || !(s.pe[h] > 0)) continue;
          double tt = s.t[h];
          if (nhs[sd] == 0) {
            tmn[sd] = tt;
            tmx[sd] = tt;
          } else {
            tmn[sd] = fmin(tmn[sd], tt);
            tmx[sd] = fmax(tmx[sd], tt);
          }
          ++nhs[sd];
          ps = ps + s.pe[h];
          pt = pt + tt * s.pe[h];
        }
        spe[sd] = ps;
        st[sd] = ps > 0 ? pt / ps : 0.0;
      }
      double slope = (tot * sxy - sx * sy) / (tot * syy - sy * sy);
      int ld = t.ldir[t.sector[E.bar[mem_e[base]]]];
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
      const int64_t slot = a + nclus;
      cl_ev[slot] = evkey;
      cl_type[slot] = type;
      cl_poscalc[slot] = calc;
      cl_sidehits[slot * 2] = nhs[0];
      cl_sidehits[slot * 2 + 1] = nhs[1];
      double *d = cl_d + slot * 12;
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
      cl_moff_e[nclus] = base;
      ++nclus;
    }
  }
  cl_moff_e[nclus] = nm;
  ncl[g] = nclus;
  nmem[g] = nm;
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
      if (!out->cl_mem_off) return set_error(c, CRV_E_HIP, "workspace allocation failed");
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
    return set_error(c, CRV_E_HIP, "workspace allocation failed");
  Ev e{bar, sipm, seq};
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  kmark(c, 3, 0, s);
  k_coinc<<<nblocks(nev, 64), 64, 0, s>>>(c->p, c->t, nev, eoff, evk, time, pe, e, sc, ncl, nmem, cl_ev, cl_type,
                                          cl_pc, cl_sh, cl_d, cl_moff, mem, c->flags);
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
