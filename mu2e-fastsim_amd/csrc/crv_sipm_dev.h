// Device pieces shared by the FP64 SiPM kernels (k_sipm over step photon streams, k_sipm_list over photon lists):
// the lane's charge records, pixel recovery, and the cross-talk cascade of a discharge (reference response/sipm.py:
// 146-177: a fired pixel proposes each existing 4-neighbour with probability q_ct; the candidates of a generation,
// in ascending pixel order, fire with 1 - (1 - p_fire(v))^m for m proposals).
#pragma once
#include "crv_internal.h"

namespace crv {

constexpr int SIPM_CMAX = 64;     // cascade candidates per generation and lane
constexpr int SIPM_MAX_GEN = 31;  // cascade generations the draw layout encodes

// LDS of a workgroup of 64 lanes: each lane's "ever fired" bitmap of its 41 x 41 pixels and its cascade lists
struct SipmLds {
  uint32_t bm[CRV_DEAD_WORDS * 64];
  uint16_t cur[SIPM_CMAX * 64], cand[SIPM_CMAX * 64];
  uint8_t mult[SIPM_CMAX * 64];
};

// a lane's charge records, in processing order, in its region [base, base + cap) of the record pool
struct LaneRecs {
  double *time, *v;
  int32_t *row;
  int16_t *pix;
  int64_t base, cap;
  int32_t cnt;
};

__device__ __forceinline__ bool px_get(const SipmLds &L, int lane, int pix) {
  return (L.bm[(pix >> 5) * 64 + lane] >> (pix & 31)) & 1u;
}
__device__ __forceinline__ void px_set(SipmLds &L, int lane, int pix) { L.bm[(pix >> 5) * 64 + lane] |= 1u << (pix & 31); }
__device__ __forceinline__ bool px_dead(const uint32_t *dead, int pix) { return (dead[pix >> 5] >> (pix & 31)) & 1u; }

// photon-map bin of a CDF point: the first bin whose cumulative content exceeds x (numpy searchsorted side="right")
__device__ __forceinline__ int upper_bound(const double *__restrict__ a, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// e^{-(t - t_last) / tau} for the pixel's last discharge (the newest record of the pixel), 0 if it never fired
__device__ __forceinline__ double recovery(const crv_params &p, const SipmLds &L, int lane, const LaneRecs &R, int pix,
                                           double t) {
  if (!px_get(L, lane, pix)) return 0.0;
  for (int32_t q = R.cnt - 1; q >= 0; --q)
    if (R.pix[R.base + q] == pix) return fm_exp(-(t - R.time[R.base + q]) / p.time_constant);
  return 0.0;
}

// append a discharge; false when the lane's region is full
__device__ __forceinline__ bool add_record(LaneRecs &R, double t, double v, int32_t row, int pix) {
  if (R.cnt >= R.cap) return false;
  R.time[R.base + R.cnt] = t;
  R.v[R.base + R.cnt] = v;
  R.row[R.base + R.cnt] = row;
  R.pix[R.base + R.cnt] = (int16_t)pix;
  ++R.cnt;
  return true;
}

// The cross-talk cascade of a discharge of pix0 at tph, already recorded, breadth-first. Draws are keyed by the
// photon (ev, ch, j, rank r), the generation and the pixel. on_fire(pix, gen, v) runs after each recorded firing.
// Returns false on overflow (region or candidate list full, or deeper than SIPM_MAX_GEN: F_CASCADE also set).
template <class OnFire>
__device__ __forceinline__ bool cascade(const crv_params &p, const Tables &t, const Derived &dv, const Draw &dr,
                                        uint32_t ev, uint32_t ch, uint32_t j, uint32_t r, double tph, int32_t row,
                                        int pix0, LaneRecs &R, SipmLds &L, int lane, int *flags, OnFire &&on_fire) {
  bool overflow = false;
  int ncur = 1, gnum = 0;
  L.cur[lane] = (uint16_t)pix0;
  while (ncur > 0) {
    int ncand = 0;
    for (int ai = 0; ai < ncur; ++ai) {
      const int f = L.cur[ai * 64 + lane];
      const int fx = f / CRV_NPIX, fy = f % CRV_NPIX;
      const U4 ux = dr(KIND_XTALK, ev, ch, j, pack_photon(r, (uint32_t)gnum, (uint32_t)f));
      const uint32_t uw[4] = {ux.x, ux.y, ux.z, ux.w};
      const int nxs[4] = {fx - 1, fx + 1, fx, fx};
      const int nys[4] = {fy, fy, fy - 1, fy + 1};
      const bool oks[4] = {fx > 0, fx + 1 < p.n_pixels_x, fy > 0, fy + 1 < p.n_pixels_y};
      for (int nb = 0; nb < 4; ++nb) {
        if (!oks[nb] || !(unif(uw[nb]) < dv.q_ct)) continue;
        const int c = nxs[nb] * CRV_NPIX + nys[nb];
        int at = 0;
        while (at < ncand && L.cand[at * 64 + lane] < c) ++at;
        if (at < ncand && L.cand[at * 64 + lane] == c) {
          L.mult[at * 64 + lane] += 1;
        } else if (ncand < SIPM_CMAX) {
          for (int q = ncand; q > at; --q) {
            L.cand[q * 64 + lane] = L.cand[(q - 1) * 64 + lane];
            L.mult[q * 64 + lane] = L.mult[(q - 1) * 64 + lane];
          }
          L.cand[at * 64 + lane] = (uint16_t)c;
          L.mult[at * 64 + lane] = 1;
          ++ncand;
        } else {
          overflow = true;
        }
      }
    }
    if (ncand == 0 || overflow) break;
    ++gnum;
    if (gnum > SIPM_MAX_GEN) {
      atomicAdd(&flags[F_CASCADE], 1);
      overflow = true;
      break;
    }
    ncur = 0;
    for (int ai = 0; ai < ncand; ++ai) {
      const int c = L.cand[ai * 64 + lane];
      const int m = L.mult[ai * 64 + lane];
      if (px_dead(t.dead, c)) continue;
      const double vc = p.overvoltage * (1.0 - recovery(p, L, lane, R, c, tph));
      const double pfc = p.p1 * (1.0 - fm_exp(-vc / p.p2));
      const double q1 = 1.0 - pfc;
      double qm = q1;
      for (int mm = 2; mm <= m; ++mm) qm = qm * q1;
      const double prc = 1.0 - qm;
      const double uc = unif(dr(KIND_XTFIRE, ev, ch, j, pack_photon(r, (uint32_t)gnum, (uint32_t)c)).x);
      if (!(uc < prc)) continue;
      if (!add_record(R, tph, vc, row, c)) {
        overflow = true;
        break;
      }
      px_set(L, lane, c);
      L.cur[ncur * 64 + lane] = (uint16_t)c;
      ++ncur;
      on_fire(c, gnum, vc);
    }
    if (overflow) break;
  }
  return !overflow;
}

}  // namespace crv
