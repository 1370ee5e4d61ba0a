// SiPM stage on a photon list (crv_sipm_run): photon arrivals plus dark noise -> SiPM charges with cross-talk and
// trap afterpulses (FP64).
//
// Follows the reference's response/sipm.py:102-198 simulate with thermal_chans, t_start and t_end:
//  * photons after t_end are dropped; each SiPM of thermal_labels gets Poisson(thermal_rate / p_fire(Vov) *
//    (t_end - t_start)) dark discharges at uniform times and uniform pixels (photon id -1), merged with its photons in
//    time order (photons first on equal times);
//  * every entry fires like a photon (recovery, p_fire, cross-talk cascade); each firing is captured by trap i with
//    probability prob_i / p_fire(Vov) and releases an afterpulse on its pixel after an Exp(lifetime_i) delay;
//  * a SiPM's afterpulses are tried after all its entries, in time order, against the pixel's last discharge at that
//    point, with p_fire(v) and no cross-talk; those after t_end are dropped.
// Draws are keyed by the entry's identity: a photon is (label, j, rank) as crv_photons_run gives it (or its list
// position and rank 0), dark discharge k of a SiPM is (label, DARK_J, k); cascade, trap and afterpulse draws add the
// generation and the pixel.
//
// Layout: entries (photons in list order, then dark discharges) are sorted by (label, time, entry) with two stable
// radix sorts, cut into SiPMs by a run-length encoding of the labels, and one lane simulates one SiPM in the record
// region its entry count sizes (grown and run again on overflow). The output kernel merges each SiPM's entry records
// with its afterpulse records by time.
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "crv_internal.h"
#include "crv_sipm_dev.h"

namespace crv {

namespace {

constexpr uint32_t DARK_J = 0xFFFFFFFFu;  // step ordinal of dark discharges (photon list positions are smaller)
constexpr int64_t NO_LABEL = INT64_MAX;   // dropped photons: sorted after every SiPM

struct Entries {  // photons (list order), then dark discharges
  int64_t *label;
  double *time;
  uint32_t *j, *r;  // draw identity
  int16_t *pix;     // dark discharges: their pixel; photons: -1 (drawn from the photon map)
  int64_t *photon;
};

struct Traps {  // afterpulse candidates of each lane, in its region of the record pool
  double *t;
  int32_t *e;  // entry whose cascade made the trapped firing
  int16_t *pix;
  uint8_t *gen, *typ;
};

struct TrapP {
  double q[2], life[2];
};

__device__ __forceinline__ void label_split(int64_t label, uint32_t &ev, uint32_t &ch, uint32_t &evhi) {
  const int64_t evkey = label / CRV_CHAN_MULT;
  ev = (uint32_t)evkey;
  evhi = (uint32_t)(evkey >> 32);
  ch = (uint32_t)(label % CRV_CHAN_MULT);
}

__global__ void k_list_entries(crv_photon_list ph, double t_end, Entries E, int *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ph.n) return;
  const int64_t lb = ph.label[i];
  const double t = ph.time[i];
  const uint32_t j = ph.j ? ph.j[i] : (uint32_t)i, r = ph.j ? ph.rank[i] : 0u;
  if (lb < 0 || !isfinite(t) || j == DARK_J || r > 0xFFFFu) atomicAdd(&flags[F_BADINPUT], 1);
  E.label[i] = t <= t_end ? lb : NO_LABEL;
  E.time[i] = t;
  E.j[i] = j;
  E.r[i] = r;
  E.pix[i] = -1;
  E.photon[i] = ph.photon[i];
}

__global__ void k_dark_count(crv_params p, const int64_t *__restrict__ tl, int64_t nt, double lam,
                             int32_t *__restrict__ nd, int *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int64_t lb = tl[i];
  nd[i] = 0;
  if (lb < 0) {
    atomicAdd(&flags[F_BADINPUT], 1);
    return;
  }
  uint32_t ev, ch, evhi;
  label_split(lb, ev, ch, evhi);
  bool capped = false;
  const long long k = poisson(lam, Draw{(uint32_t)p.seed, evhi}, KIND_THERMAL_N, ev, ch, 0u, &capped);
  if (capped) atomicAdd(&flags[F_CAPPED], 1);
  if (k > 0xFFFF) {  // the rank field of the draw layout holds 16 bits
    atomicAdd(&flags[F_COUNTER], 1);
    return;
  }
  nd[i] = (int32_t)k;
}

__global__ void k_dark_fill(crv_params p, const int64_t *__restrict__ tl, int64_t nt, const int32_t *__restrict__ nd,
                            const int64_t *__restrict__ doff, int64_t n0, double t_start, double t_end, Entries E) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const int64_t lb = tl[i];
  uint32_t ev, ch, evhi;
  label_split(lb, ev, ch, evhi);
  const Draw dr{(uint32_t)p.seed, evhi};
  for (int32_t k = 0; k < nd[i]; ++k) {
    const int64_t e = n0 + doff[i] + k;
    const U4 u = dr(KIND_THERMAL, ev, ch, (uint32_t)k, 0u);
    E.label[e] = lb;
    E.time[e] = t_start + (t_end - t_start) * unif(u.x);
    const int px = (int)floor(unif(u.y) * (double)p.n_pixels_x), py = (int)floor(unif(u.z) * (double)p.n_pixels_y);
    E.pix[e] = (int16_t)(px * CRV_NPIX + py);
    E.j[e] = DARK_J;
    E.r[e] = (uint32_t)k;
    E.photon[e] = -1;
  }
}

// order-preserving key of a time (-0 as +0, so that equal times tie and keep their entry order)
__global__ void k_time_keys(const double *__restrict__ time, int64_t n, uint64_t *__restrict__ key,
                            int32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = time[i] == 0.0 ? 0.0 : time[i];
  const uint64_t b = (uint64_t)__double_as_longlong(t);
  key[i] = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  idx[i] = (int32_t)i;
}

__global__ void k_label_keys(const int32_t *__restrict__ idx, const int64_t *__restrict__ label, int64_t n,
                             uint64_t *__restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (uint64_t)label[idx[i]];
}

__global__ void k_list_caps(const int32_t *__restrict__ cnt, int64_t S, int64_t *__restrict__ cap, int factor) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S) return;
  const int64_t v = cnt[g];
  cap[g] = v + (v * factor) / 4 + 8;
}

// one lane per SiPM: its entries in time order, then its afterpulses
__global__ __launch_bounds__(64) void k_sipm_list(crv_params p, Tables t, Derived dv, TrapP tp, double t_end,
                                                  int64_t S, const uint64_t *__restrict__ ulab,
                                                  const int64_t *__restrict__ eoff, const int32_t *__restrict__ order,
                                                  Entries E, const int64_t *__restrict__ coff, double *__restrict__ c_time,
                                                  double *__restrict__ c_v, int32_t *__restrict__ c_row,
                                                  int16_t *__restrict__ c_pix, Traps ap, int32_t *__restrict__ c_count,
                                                  int32_t *__restrict__ c_count1, int *__restrict__ flags) {
  __shared__ SipmLds L;
  const int lane = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * 64 + lane;
  if (g >= S) return;
  uint32_t ev, ch, evhi;
  label_split((int64_t)ulab[g], ev, ch, evhi);
  const Draw dr{(uint32_t)p.seed, evhi};
  LaneRecs R{c_time, c_v, c_row, c_pix, coff[g], coff[g + 1] - coff[g], 0};
  for (int w = 0; w < CRV_DEAD_WORDS; ++w) L.bm[w * 64 + lane] = 0u;
  const bool traps = tp.q[0] > 0.0 || tp.q[1] > 0.0;
  int32_t nap = 0;
  bool overflow = false;
  for (int64_t q = eoff[g]; q < eoff[g + 1] && !overflow; ++q) {
    const int32_t e = order[q];
    const double tph = E.time[e];
    const uint32_t j = E.j[e], r = E.r[e];
    const U4 u = dr(KIND_PIXEL, ev, ch, j, pack_photon(r, 0u, 0u));
    int pix = E.pix[e];
    if (pix < 0) {
      const int bin = upper_bound(t.cdf, t.ncdf, unif(u.x) * t.cdf_total);
      int px = bin / t.map_ny, py = bin % t.map_ny;
      if (p.pixel_lrint) {
        px += unif(u.y) >= 0.5 ? 1 : 0;
        py += unif(u.z) >= 0.5 ? 1 : 0;
      }
      pix = px * CRV_NPIX + py;
    }
    if (px_dead(t.dead, pix)) continue;
    const double v = p.overvoltage * (1.0 - recovery(p, L, lane, R, pix, tph));
    const double pf = p.p1 * (1.0 - fm_exp(-v / p.p2));
    if (!(unif(u.w) < 1.0 - (1.0 - pf))) continue;
    // trap captures of a firing (KIND_TRAP: trap 0 test and lifetime, trap 1 test and lifetime)
    auto on_fire = [&](int fpix, int gen, double) {
      if (!traps) return;
      const U4 ut = dr(KIND_TRAP, ev, ch, j, pack_photon(r, (uint32_t)gen, (uint32_t)fpix));
      const uint32_t w[4] = {ut.x, ut.y, ut.z, ut.w};
      for (int ty = 0; ty < 2; ++ty) {
        if (!(unif(w[2 * ty]) < tp.q[ty])) continue;
        if (nap >= R.cap) {
          overflow = true;
          return;
        }
        const int64_t at = R.base + nap++;
        ap.t[at] = tph - tp.life[ty] * fm_log(unif(w[2 * ty + 1]));
        ap.e[at] = e;
        ap.pix[at] = (int16_t)fpix;
        ap.gen[at] = (uint8_t)gen;
        ap.typ[at] = (uint8_t)ty;
      }
    };
    if (!add_record(R, tph, v, e, pix)) {
      overflow = true;
      break;
    }
    px_set(L, lane, pix);
    on_fire(pix, 0, v);
    if (!cascade(p, t, dv, dr, ev, ch, j, r, tph, e, pix, R, L, lane, flags, on_fire)) overflow = true;
  }
  const int32_t n1 = R.cnt;
  if (!overflow && nap) {
    // afterpulses in time order (insertion sort, capture order on equal times), then tried one after the other
    for (int32_t a = 1; a < nap; ++a) {
      const int64_t x = R.base + a;
      const double ta = ap.t[x];
      const int32_t ea = ap.e[x];
      const int16_t pa = ap.pix[x];
      const uint8_t ga = ap.gen[x], ya = ap.typ[x];
      int64_t y = x - 1;
      while (y >= R.base && ap.t[y] > ta) {
        ap.t[y + 1] = ap.t[y];
        ap.e[y + 1] = ap.e[y];
        ap.pix[y + 1] = ap.pix[y];
        ap.gen[y + 1] = ap.gen[y];
        ap.typ[y + 1] = ap.typ[y];
        --y;
      }
      ap.t[y + 1] = ta;
      ap.e[y + 1] = ea;
      ap.pix[y + 1] = pa;
      ap.gen[y + 1] = ga;
      ap.typ[y + 1] = ya;
    }
    for (int32_t a = 0; a < nap; ++a) {
      const int64_t x = R.base + a;
      const double t1 = ap.t[x];
      if (t1 > t_end) continue;
      const int pix = ap.pix[x];
      const int32_t e = ap.e[x];
      const double v = p.overvoltage * (1.0 - recovery(p, L, lane, R, pix, t1));
      const double pf = p.p1 * (1.0 - fm_exp(-v / p.p2));
      const U4 ua = dr(KIND_AFTER, ev, ch, E.j[e], pack_photon(E.r[e], (uint32_t)ap.gen[x], (uint32_t)pix));
      if (!(unif(ap.typ[x] ? ua.y : ua.x) < pf)) continue;
      if (!add_record(R, t1, v, e, pix)) {
        overflow = true;
        break;
      }
    }
  }
  if (overflow) atomicAdd(&flags[F_OVERFLOW], 1);
  c_count[g] = R.cnt;
  c_count1[g] = n1;
}

// each SiPM's entry records and afterpulse records merged by time (entry records first on equal times)
__global__ void k_sipm_list_out(crv_params p, int64_t S, const uint64_t *__restrict__ ulab,
                                const int64_t *__restrict__ coff, const int32_t *__restrict__ c_count,
                                const int32_t *__restrict__ c_count1, const int64_t *__restrict__ ooff,
                                const double *__restrict__ c_time, const double *__restrict__ c_v,
                                const int32_t *__restrict__ c_row, const int64_t *__restrict__ e_photon,
                                int64_t *__restrict__ o_label, double *__restrict__ o_time,
                                double *__restrict__ o_charge, double *__restrict__ o_pe, int64_t *__restrict__ o_photon) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S) return;
  const int64_t base = coff[g], o = ooff[g];
  const int32_t n = c_count[g], n1 = c_count1[g];
  int32_t a = 0, b = n1;
  for (int32_t k = 0; k < n; ++k) {
    const bool from_a = b >= n || (a < n1 && c_time[base + a] <= c_time[base + b]);
    const int64_t q = base + (from_a ? a++ : b++);
    const double v = c_v[q];
    o_label[o + k] = (int64_t)ulab[g];
    o_time[o + k] = c_time[q];
    o_charge[o + k] = p.capacitance * v;
    o_pe[o + k] = v / p.overvoltage;
    o_photon[o + k] = e_photon[c_row[q]];
  }
}

}  // namespace

int sipm_list_stage(Ctx *c, const crv_photon_list *ph, const crv_sipm_noise *nz, crv_charges *out, hipStream_t s) {
  const crv_params &p = c->p;
  const int64_t n = ph->n, nt = nz->n_thermal;
  if (n < 0 || nt < 0) return set_error(c, CRV_E_ARG, "sipm: negative photon or thermal-label count");
  if (n && (!ph->label || !ph->time || !ph->photon)) return set_error(c, CRV_E_ARG, "sipm: NULL photon array");
  if (n && !ph->j != !ph->rank) return set_error(c, CRV_E_ARG, "sipm: photon j and rank go together");
  if (nt && !nz->thermal_labels) return set_error(c, CRV_E_ARG, "sipm: NULL thermal_labels");
  auto finite_nonneg = [](double x) { return std::isfinite(x) && x >= 0.0; };
  if (!std::isfinite(nz->t_end)) return set_error(c, CRV_E_ARG, "sipm: t_end must be finite");
  if (!finite_nonneg(nz->thermal_rate)) return set_error(c, CRV_E_ARG, "sipm: thermal_rate must be finite and >= 0");
  if (!finite_nonneg(nz->trap0_prob) || !finite_nonneg(nz->trap1_prob))
    return set_error(c, CRV_E_ARG, "sipm: trap0_prob and trap1_prob must be finite and >= 0");
  if ((nz->trap0_prob > 0 && !(std::isfinite(nz->trap0_lifetime) && nz->trap0_lifetime > 0)) ||
      (nz->trap1_prob > 0 && !(std::isfinite(nz->trap1_lifetime) && nz->trap1_lifetime > 0)))
    return set_error(c, CRV_E_ARG, "sipm: trap lifetimes must be finite and positive");
  const bool thermal = nz->thermal_rate > 0 && nt > 0;
  if (thermal && !(std::isfinite(nz->t_start) && nz->t_start <= nz->t_end))
    return set_error(c, CRV_E_ARG, "sipm: t_start must be finite and at most t_end");
  const double pf_full = p.p1 * (1.0 - fm_exp(-p.overvoltage / p.p2));
  if (!(pf_full > 0.0) && (thermal || nz->trap0_prob > 0 || nz->trap1_prob > 0))
    return set_error(c, CRV_E_ARG, "sipm: dark noise and traps need a positive firing probability at Vov");
  out->n_photons = 0;
  out->capped = 0;
  int rc = 0;
  CRV_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int) * F_NFLAGS, s));
  // dark discharges per thermal label
  int64_t D = 0;
  int32_t *nd = nullptr;
  int64_t *doff = nullptr;
  if (thermal) {
    nd = (int32_t *)ws(c, 606, nt * 4 + 8);
    doff = (int64_t *)ws(c, 607, (nt + 1) * 8);
    if (!nd || !doff) return ws_fail(c);
    const double lam = nz->thermal_rate / pf_full * (nz->t_end - nz->t_start);
    k_dark_count<<<nblocks(nt, 256), 256, 0, s>>>(p, nz->thermal_labels, nt, lam, nd, c->flags);
    CRV_HIP(c, hipGetLastError());
    if ((rc = scan_i32_to_i64(c, nd, doff, nt, 608, s))) return rc;
    if ((rc = read_i64(c, doff + nt, &D, s))) return rc;
  }
  const int64_t E = n + D;
  if (E >= INT32_MAX) return set_error(c, CRV_E_ARG, "sipm: more than 2^31 - 1 photons and dark discharges");
  if (E == 0) {
    out->n = 0;
    return CRV_OK;
  }
  Entries en{(int64_t *)ws(c, 600, E * 8), (double *)ws(c, 601, E * 8), (uint32_t *)ws(c, 602, E * 4),
             (uint32_t *)ws(c, 603, E * 4), (int16_t *)ws(c, 604, E * 2), (int64_t *)ws(c, 605, E * 8)};
  if (!en.label || !en.time || !en.j || !en.r || !en.pix || !en.photon) return ws_fail(c);
  if (n) k_list_entries<<<nblocks(n, 256), 256, 0, s>>>(*ph, nz->t_end, en, c->flags);
  if (D) k_dark_fill<<<nblocks(nt, 256), 256, 0, s>>>(p, nz->thermal_labels, nt, nd, doff, n, nz->t_start, nz->t_end, en);
  CRV_HIP(c, hipGetLastError());
  // (label, time, entry) order: a stable sort by time, then a stable sort by label
  uint64_t *k1 = (uint64_t *)ws(c, 609, E * 8), *k2 = (uint64_t *)ws(c, 611, E * 8);
  int32_t *i1 = (int32_t *)ws(c, 610, E * 4), *i2 = (int32_t *)ws(c, 612, E * 4), *order = (int32_t *)ws(c, 616, E * 4);
  if (!k1 || !k2 || !i1 || !i2 || !order) return ws_fail(c);
  k_time_keys<<<nblocks(E, 256), 256, 0, s>>>(en.time, E, k1, i1);
  size_t tb = 0;
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k1, k2, i1, i2, (int)E, 0, 64, s));
  void *tmp = ws(c, 613, tb);
  if (!tmp) return ws_fail(c);
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k1, k2, i1, i2, (int)E, 0, 64, s));
  k_label_keys<<<nblocks(E, 256), 256, 0, s>>>(i2, en.label, E, k1);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k1, k2, i2, order, (int)E, 0, 64, s));
  // SiPMs: runs of equal labels (the run of dropped photons last)
  uint64_t *ulab = (uint64_t *)ws(c, 617, E * 8);
  int32_t *ucnt = (int32_t *)ws(c, 618, E * 4 + 4);
  int64_t *nruns = (int64_t *)ws(c, 619, 16);
  if (!ulab || !ucnt || !nruns) return ws_fail(c);
  size_t tr = 0;
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(nullptr, tr, k2, ulab, ucnt, nruns, (int)E, s));
  void *tmp2 = ws(c, 620, tr);
  if (!tmp2) return ws_fail(c);
  CRV_HIP(c, hipcub::DeviceRunLengthEncode::Encode(tmp2, tr, k2, ulab, ucnt, nruns, (int)E, s));
  int64_t nr = 0;
  if ((rc = read_i64(c, nruns, &nr, s))) return rc;
  uint64_t lastlab = 0;
  CRV_HIP(c, hipMemcpyAsync(&lastlab, ulab + nr - 1, 8, hipMemcpyDeviceToHost, s));
  int hf[F_NFLAGS];
  CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
  CRV_HIP(c, hipStreamSynchronize(s));
  if (hf[F_BADINPUT])
    return set_error(c, CRV_E_ARG, "sipm: negative label, non-finite photon time, or a rank above 65535");
  if (hf[F_COUNTER]) return set_error(c, CRV_E_OVERFLOW, "sipm: more than 65535 dark discharges in one SiPM");
  out->capped = hf[F_CAPPED];
  const int64_t S = nr - (lastlab == (uint64_t)NO_LABEL ? 1 : 0);
  int64_t *eoff = (int64_t *)ws(c, 621, (S + 1) * 8);
  if (!eoff) return ws_fail(c);
  if ((rc = scan_i32_to_i64(c, ucnt, eoff, S, 622, s))) return rc;
  if ((rc = read_i64(c, eoff + S, &out->n_photons, s))) return rc;
  if (S == 0) {
    out->n = 0;
    return CRV_OK;
  }
  const bool traps = nz->trap0_prob > 0 || nz->trap1_prob > 0;
  const TrapP tp{{nz->trap0_prob / pf_full, nz->trap1_prob / pf_full}, {nz->trap0_lifetime, nz->trap1_lifetime}};
  int64_t *cap = (int64_t *)ws(c, 623, (S + 1) * 8), *coff = (int64_t *)ws(c, 624, (S + 1) * 8);
  int32_t *cc = (int32_t *)ws(c, 625, S * 4 + 4), *cc1 = (int32_t *)ws(c, 626, S * 4 + 4);
  int64_t *ooff = (int64_t *)ws(c, 627, (S + 1) * 8);
  if (!cap || !coff || !cc || !cc1 || !ooff) return ws_fail(c);
  for (int factor = 2; factor <= 64; factor *= 4) {
    k_list_caps<<<nblocks(S, 256), 256, 0, s>>>(ucnt, S, cap, factor);
    if ((rc = scan_i64(c, cap, coff, S, 628, s))) return rc;
    int64_t total = 0;
    if ((rc = read_i64(c, coff + S, &total, s))) return rc;
    double *c_time = (double *)ws(c, 629, total * 8 + 8), *c_v = (double *)ws(c, 630, total * 8 + 8);
    int32_t *c_row = (int32_t *)ws(c, 631, total * 4 + 8);
    int16_t *c_pix = (int16_t *)ws(c, 632, total * 2 + 8);
    if (!c_time || !c_v || !c_row || !c_pix) return ws_fail(c);
    Traps ap{};
    if (traps) {
      ap = Traps{(double *)ws(c, 633, total * 8 + 8), (int32_t *)ws(c, 634, total * 4 + 8),
                 (int16_t *)ws(c, 635, total * 2 + 8), (uint8_t *)ws(c, 636, total + 8),
                 (uint8_t *)ws(c, 637, total + 8)};
      if (!ap.t || !ap.e || !ap.pix || !ap.gen || !ap.typ) return ws_fail(c);
    }
    CRV_HIP(c, hipMemsetAsync(c->flags + F_OVERFLOW, 0, sizeof(int) * 2, s));
    k_sipm_list<<<nblocks(S, 64), 64, 0, s>>>(p, c->t, c->d, tp, nz->t_end, S, ulab, eoff, order, en, coff, c_time,
                                              c_v, c_row, c_pix, ap, cc, cc1, c->flags);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    CRV_HIP(c, hipStreamSynchronize(s));
    if (hf[F_CASCADE]) return set_error(c, CRV_E_OVERFLOW, "sipm: cross-talk cascade deeper than 31 generations");
    if (hf[F_OVERFLOW]) continue;  // more room per SiPM, then again
    if ((rc = scan_i32_to_i64(c, cc, ooff, S, 638, s))) return rc;
    int64_t nq = 0;
    if ((rc = read_i64(c, ooff + S, &nq, s))) return rc;
    out->n = nq;
    if (nq > out->capacity) return set_error(c, CRV_E_CAPACITY, "sipm: output capacity too small");
    k_sipm_list_out<<<nblocks(S, 256), 256, 0, s>>>(p, S, ulab, coff, cc, cc1, ooff, c_time, c_v, c_row, en.photon,
                                                    out->label, out->time, out->charge, out->pe, out->photon);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipStreamSynchronize(s));
    return CRV_OK;
  }
  return set_error(c, CRV_E_OVERFLOW, "sipm: charges exceed 17 per entry or a cascade generation exceeds 64 pixels");
}

}  // namespace crv
