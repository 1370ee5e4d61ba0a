// Portable double-precision math, Philox4x32-10 and the samplers of the CRV response chain.
//
// Every function is a fixed sequence of IEEE-754 double operations (the library is built with
// -ffp-contract=off), so a count, a delay or a noise sample is a pure function of the seed and the draw layout:
//  * fm_exp: Cody & Waite (1980) reduction x = k ln2 + r (ln2 split in 32-bit parts), Taylor series of e^r to
//    degree 13, scaled by 2^k.
//  * fm_log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) to s^23.
//  * fm_sin / fm_cos: reduction by pi/2 in three 30-bit parts, Taylor series on |r| <= pi/4, quadrant swap.
//  * fm_asin: Maclaurin series to z^24 on |x| <= 1/2, asin x = pi/2 - 2 asin(sqrt((1-x)/2)) above.
//  * Philox4x32-10: Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
//  * Poisson: inversion (mu < 10); PTRS, W. Hoermann, Insurance: Mathematics and Economics 12 (1993) 39-45.
//  * Normal: Box-Muller.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "fmath_consts.h"

#define CRV_HD __host__ __device__ __forceinline__

namespace crv {

template <int N>
CRV_HD double horner(const double (&c)[N], double z) {
  double p = c[N - 1];
#pragma unroll
  for (int i = N - 2; i >= 0; --i) p = p * z + c[i];
  return p;
}

struct Coef {
  static constexpr double exp_c[FM_EXP_C_N] = FM_EXP_C_INIT;
  static constexpr double log_c[FM_LOG_C_N] = FM_LOG_C_INIT;
  static constexpr double sin_c[FM_SIN_C_N] = FM_SIN_C_INIT;
  static constexpr double cos_c[FM_COS_C_N] = FM_COS_C_INIT;
  static constexpr double asin_c[FM_ASIN_C_N] = FM_ASIN_C_INIT;
  static constexpr double stirling[FM_STIRLING_N] = FM_STIRLING_INIT;
  static constexpr double fact[FM_FACT_N] = FM_FACT_INIT;
};

CRV_HD double fm_exp(double x) {
  if (x != x) return x;
  if (x > FM_EXP_HI) return INFINITY;
  if (x < FM_EXP_LO) return 0.0;
  double k = rint(x * FM_INV_LN2);
  double r = (x - k * FM_LN2_HI) - k * FM_LN2_LO;
  return ldexp(horner(Coef::exp_c, r), (int)k);
}

CRV_HD double fm_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int e;
  double m = frexp(x, &e);
  if (m < FM_SQRT_HALF) {
    m = m * 2.0;
    e = e - 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double p = Coef::log_c[FM_LOG_C_N - 1];
#pragma unroll
  for (int i = FM_LOG_C_N - 2; i >= 1; --i) p = p * z + Coef::log_c[i];
  double q = z * p;
  double lm = (s + s * q) * 2.0;
  double de = (double)e;
  return de * FM_LN2_HI + (lm + de * FM_LN2_LO);
}

CRV_HD void fm_reduce(double x, double &r, int &q) {
  double k = rint(x * FM_TWO_OVER_PI);
  r = ((x - k * FM_PIO2_1) - k * FM_PIO2_2) - k * FM_PIO2_3;
  q = (int)(((long long)k) & 3);
}

CRV_HD double fm_sinp(double r) { return r * horner(Coef::sin_c, r * r); }
CRV_HD double fm_cosp(double r) { return horner(Coef::cos_c, r * r); }

CRV_HD double fm_sin(double x) {
  double r;
  int q;
  fm_reduce(x, r, q);
  if (q == 0) return fm_sinp(r);
  if (q == 1) return fm_cosp(r);
  if (q == 2) return -fm_sinp(r);
  return -fm_cosp(r);
}

CRV_HD double fm_cos(double x) {
  double r;
  int q;
  fm_reduce(x, r, q);
  if (q == 0) return fm_cosp(r);
  if (q == 1) return -fm_sinp(r);
  if (q == 2) return -fm_cosp(r);
  return fm_sinp(r);
}

CRV_HD double fm_asin(double x) {
  double a = fabs(x);
  if (a > 1.0) return NAN;
  double r;
  if (a <= 0.5) {
    r = a * horner(Coef::asin_c, a * a);
  } else {
    double w = (1.0 - a) * 0.5;
    double s = sqrt(w);
    r = FM_PI_2 - 2.0 * (s * horner(Coef::asin_c, w));
  }
  return x < 0 ? -r : r;
}

// ---------------------------------------------------------------------------------------------------------------
// Philox4x32-10 and the draw layout: key = (seed, kind + 16 * evkey_hi), counter = (evkey_lo, c1, c2, c3).

enum Kind : uint32_t {
  KIND_COUNT = 1, KIND_DELAY = 2, KIND_PIXEL = 3, KIND_XTALK = 4, KIND_XTFIRE = 5, KIND_TDC0 = 6, KIND_NOISE = 7,
  KIND_THERMAL_N = 8, KIND_THERMAL = 9, KIND_TRAP = 10, KIND_AFTER = 11  // dark noise and trap afterpulses
};

struct U4 {
  uint32_t x, y, z, w;
};

CRV_HD U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32 x 32 -> 64 multiply per word gives both halves (a single v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
  }
  return c;
}

CRV_HD double unif(uint32_t x) { return ((double)x + 0.5) * 0x1p-32; }

struct Draw {
  uint32_t seed, evhi;
  CRV_HD U4 operator()(uint32_t kind, uint32_t ev, uint32_t c1, uint32_t c2, uint32_t c3) const {
    return philox(U4{ev, c1, c2, c3}, seed, kind + 16u * evhi);
  }
};

CRV_HD uint32_t pack_photon(uint32_t rank, uint32_t gen, uint32_t pix) { return (rank << 16) | (gen << 11) | pix; }

// ---------------------------------------------------------------------------------------------------------------
// samplers

CRV_HD double loggam_int(double x) {
  if (x < 10.0) return fm_log(Coef::fact[(int)x - 1]);
  double z = 1.0 / x;
  double z2 = z * z;
  double s = z * horner(Coef::stirling, z2);
  double t1 = (x - 0.5) * fm_log(x);
  return ((t1 - x) + FM_LN_SQRT_2PI) + s;
}

// Poisson(mu) on the draws (kind, ev, c1, c2, call); returns k, sets *capped when the rejection bound is reached.
CRV_HD long long poisson(double mu, const Draw &dr, uint32_t kind, uint32_t ev, uint32_t c1, uint32_t c2,
                         bool *capped) {
  if (mu < 10.0) {
    double u = unif(dr(kind, ev, c1, c2, 0).x);
    double p = fm_exp(-mu);
    double cdf = p;
    long long k = 0;
    while (u > cdf && k < 200) {
      k += 1;
      p = (p * mu) / (double)k;
      cdf = cdf + p;
    }
    return k;
  }
  double slam = sqrt(mu);
  double loglam = fm_log(mu);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2.0);
  U4 cache = dr(kind, ev, c1, c2, 0);
  for (int trial = 0; trial < 64; ++trial) {
    if (trial && (trial & 1) == 0) cache = dr(kind, ev, c1, c2, (uint32_t)(trial >> 1));
    double U = unif((trial & 1) ? cache.z : cache.x) - 0.5;
    double V = unif((trial & 1) ? cache.w : cache.y);
    double us = 0.5 - fabs(U);
    double kf = floor((((2.0 * a) / us + b) * U + mu) + 0.43);
    if (us >= 0.07 && V <= vr) return (long long)kf;
    if (kf < 0 || (us < 0.013 && V > us)) continue;
    double lhs = fm_log((V * invalpha) / (a / (us * us) + b));
    double rhs = (-mu + kf * loglam) - loggam_int(kf + 1.0);
    if (lhs <= rhs) return (long long)kf;
  }
  *capped = true;
  return (long long)rint(mu);
}

CRV_HD double box_muller(double u1, double u2) { return sqrt(-2.0 * fm_log(u1)) * fm_cos(FM_TWO_PI * u2); }

}  // namespace crv
