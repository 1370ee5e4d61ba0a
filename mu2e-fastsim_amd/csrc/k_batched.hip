// The FP32 batched chain: steps -> pulses in four persistent / data-parallel kernels, then the coincidence finder.
//
//  kb_sipm  one lane per SiPM: the photon streams of the counter's steps (ordered exponential delays) -> pixel,
//           recovery, firing, cross-talk cascade (reference response/sipm.py:141-198) -> the SiPM's fired-pixel log
//           (time, overvoltage), and its charge clusters as they are logged (gaps of more than the template's
//           length plus four periods) -> a descriptor per cluster that can reach the zero-suppression threshold;
//           no photon or charge table.
//  kb_wrec  one lane per cluster descriptor: a 16-sample register window of template sums, noise, ADC words and zero
//           suppression (reference reco/waveform.py:96-172) -> the cluster's ADC record (record pool) and its reco
//           waveforms, the runs of continuing digis (run pool).
//  kb_fit   one lane per reco waveform: peak finding and the Gumbel fit (reference reco/pulsefit.py:236-316).
// The SiPM kernels take tiles of SiPMs from an atomic work queue (persistent grid), optionally longest first (photon
// count), so that the lanes of a wave carry similar work. A run map (SiPM, run number) -> pool slot keeps every
// output in (event, channel, time) order whatever order the lanes fill the pools in.
//
// Arithmetic is FP32 with the hardware exp / log / sin / cos; absolute times are FP64 (photon arrival = step time +
// propagation, window starts, pulse times = TDC grid + fitted offset) and the 3 x 3 normal equations of the fit are
// solved in FP64. The draws are the Philox streams of the FP64 chain (same keys and counters), so both chains
// simulate the same photons up to rounding; their tables agree statistically, not bit for bit.
#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "crv_internal.h"

namespace crv {

namespace {

constexpr int CMAXB = 16;   // cascade candidates per generation and lane
constexpr int MAXGEN = 31;  // cascade generations
// bits of the 7 x 7 cascade window (bit (dx + 3) * 7 + dy + 3) with |dx| = 3 or |dy| = 3
constexpr uint64_t WIN_BORDER = 0x1FE0C183060FFull;
constexpr int RINGB = 16;   // open samples of the waveform window
constexpr int PSLOTS = 4;   // pulse slots per reco waveform (more: the fits run again with more)
constexpr int DWORDS = 6;   // digi-start mask words: window_samples <= 192
constexpr float PERIODF = 12.5f;
constexpr double PERIODD = 12.5;
constexpr int F_PSLOT = 7;  // a reco waveform produced more pulses than its slots
constexpr int TMPL_LDS = 1024;  // template samples staged in LDS by the waveform kernel
constexpr int FWIN = 16;        // fit-window samples staged in LDS per lane by the fit kernel
// the call's counters (work queue, record / run / descriptor / used-record counts) sit one 128-byte line apart:
// atomics on counters that share a line serialise in one L2 channel
constexpr int CL = 32;
constexpr int C_QUEUE = 0, C_REC = 2 * CL, C_RUN = 3 * CL, C_DESC = 4 * CL, C_USED = 5 * CL, C_DRANGE = 6 * CL;
constexpr int C_INTS = 7 * CL;

__device__ __forceinline__ float u32f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * 0x1p-24f; }

// Philox4x32-10 exactly as philox() of crv_math.h, with the round keys formed by scalar adds inside every call. Left
// to itself the compiler hoists the 20 round keys of each draw kind out of the photon loop, and the scalar registers
// they take are spilled and reloaded in the loop.
__device__ __forceinline__ U4 philox_rk(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
      asm volatile("" : "+s"(k0), "+s"(k1));  // keys are wave-uniform: kept in scalar registers, not hoisted
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = U4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
  }
  return c;
}

// Draw with philox_rk (the same values); seed and evhi must be wave-uniform
struct DrawR {
  uint32_t seed, evhi;
  __device__ __forceinline__ U4 operator()(uint32_t kind, uint32_t ev, uint32_t c1, uint32_t c2, uint32_t c3) const {
    return philox_rk(U4{ev, c1, c2, c3}, seed, kind + 16u * evhi);
  }
};

__device__ __forceinline__ float lgam_int_f(float x) {  // log((x - 1)!) for integer x >= 1
  if (x < 10.f) {
    float f = 1.f;
    for (int i = 2; i < (int)x; ++i) f *= (float)i;
    return __logf(f);
  }
  const float z = 1.f / x, z2 = z * z;
  const float s = z * (0.0833333333f + z2 * (-0.00277777778f + z2 * (0.000793650794f + z2 * -0.000595238095f)));
  return (x - 0.5f) * __logf(x) - x + 0.918938533f + s;
}

// Poisson(mu) on the KIND_COUNT draws of the FP64 chain: inversion below 10, PTRS (Hoermann 1993) above
__device__ __forceinline__ int poisson_f(float mu, const Draw &dr, uint32_t ev, uint32_t ch, uint32_t j, int *capped) {
  if (mu < 10.f) {
    const float u = u32f(dr(KIND_COUNT, ev, ch, j, 0).x);
    float p = __expf(-mu), cdf = p;
    int k = 0;
    while (u > cdf && k < 200) {
      k += 1;
      p = p * mu / (float)k;
      cdf += p;
    }
    return k;
  }
  const float slam = sqrtf(mu), loglam = __logf(mu);
  const float b = 0.931f + 2.53f * slam, a = -0.059f + 0.02483f * b;
  const float invalpha = 1.1239f + 1.1328f / (b - 3.4f), vr = 0.9277f - 3.6224f / (b - 2.f);
  U4 cache = dr(KIND_COUNT, ev, ch, j, 0);
  for (int trial = 0; trial < 64; ++trial) {
    if (trial && (trial & 1) == 0) cache = dr(KIND_COUNT, ev, ch, j, (uint32_t)(trial >> 1));
    const float U = u32f((trial & 1) ? cache.z : cache.x) - 0.5f;
    const float V = u32f((trial & 1) ? cache.w : cache.y);
    const float us = 0.5f - fabsf(U);
    const float kf = floorf((2.f * a / us + b) * U + mu + 0.43f);
    if (us >= 0.07f && V <= vr) return (int)kf;
    if (kf < 0.f || (us < 0.013f && V > us)) continue;
    if (__logf(V * invalpha / (a / (us * us) + b)) <= -mu + kf * loglam - lgam_int_f(kf + 1.f)) return (int)kf;
  }
  atomicAdd(capped, 1);
  return (int)rintf(mu);
}

struct BStep {  // the steps of the batch (device SoA)
  const double *edep, *z, *t;
};

// key bits of a work-queue order (crv_batch_opts.sort_tasks 1, 2, 3): above the bucket, one bit that puts the SiPMs
// of counters with several steps first; the bucket is the photon count (31 bits), its octave (at most 32) or its
// quarter octave (at most 128)
__host__ __device__ constexpr int sort_key_bits(int mode) { return mode == 1 ? 32 : mode == 2 ? 7 : 9; }

// photoelectrons per (step, SiPM), and per SiPM in total
__global__ void kb_count(crv_params p, Tables t, Derived dv, BStep st, const int32_t *__restrict__ Tdev,
                         const uint64_t *__restrict__ task_key, const int32_t *__restrict__ task_off,
                         const int32_t *__restrict__ perm, const int32_t *__restrict__ evfirst,
                         int32_t *__restrict__ nph, int64_t *__restrict__ sipm_nph, uint32_t *__restrict__ sort_key,
                         int32_t *__restrict__ sort_idx, int sort_mode, int *__restrict__ flags) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t L4 = (int64_t)(*Tdev) * 4;
  if (g >= L4) return;
  const int64_t tk = g >> 2;
  const int s = (int)(g & 3);
  const uint64_t key = task_key[tk];
  const uint32_t ev = (uint32_t)(key >> 13);
  const int ch = (int)(key & 8191u) * 4 + s;
  const int a = task_off[tk], b = task_off[tk + 1];
  const bool live = ch < t.n_channels && (t.status[ch] & 0b101) == 0;
  if (ch >= t.n_channels) atomicAdd(&flags[F_BADCHAN], 1);  // the host turns it into an error, as the FP64 chain
  const Draw dr{(uint32_t)p.seed, p.evkey_hi};
  const float ys = live ? (float)(dv.yscale * t.dev[ch]) : 0.f;
  const float H = (float)p.half_length;
  const float ws_ = (float)p.w_short, wl = (float)p.w_long, ils = 1.f / (float)p.l_short, ill = 1.f / (float)p.l_long;
  int64_t total = 0;
  for (int k = a; k < b; ++k) {
    int n = 0;
    if (live) {
      const int32_t i = perm[k];
      const uint32_t j = (uint32_t)(i - evfirst[i]);
      const float z = (float)st.z[i];
      const float d = (s & 1) == 0 ? z + H : H - z;
      const float att = ws_ * __expf(-d * ils) + wl * __expf(-d * ill);
      const float mu = (float)st.edep[i] * ys * att;
      n = mu > 0.f ? poisson_f(mu, dr, ev, (uint32_t)ch, j, &flags[F_CAPPED]) : 0;
      if (n > 65535) {  // the photon rank is 16 bits of the Philox counter
        atomicAdd(&flags[F_COUNTER], 1);
        n = 65535;
      }
    }
    nph[(int64_t)k * 4 + s] = n;
    total += n;
  }
  sipm_nph[g] = total;
  // longest first, ties in SiPM order (a stable radix sort of the complement): mode 1 by photon count, modes 2 and 3
  // by its octave or quarter octave (fewer key bits, and the SiPMs of one bucket keep their order, so that
  // neighbouring SiPMs, which share steps, stay neighbours in the queue). Above that, the SiPMs of counters with
  // several steps come first: kb_sipm merges their photon streams in memory and those with one step in registers, so
  // a wavefront of one kind runs one of the two codes instead of both.
  const uint32_t single = b - a > 1 ? 0u : 1u;
  const uint32_t v = (uint32_t)(total < 0x7FFFFFFFll ? total : 0x7FFFFFFFll);
  if (sort_mode == 1) {
    sort_key[g] = (single << 31) | (0x7FFFFFFFu - v);
  } else {
    const int mb = sort_mode == 3 ? 2 : 0;  // bits of the count below its leading one that the bucket keeps
    uint32_t bucket = 0;                    // 0 for no photons, else 1 + (octave << mb) + those bits
    if (v) {
      const int e = 31 - __clz(v);
      const uint32_t m = (e >= mb ? v >> (e - mb) : v << (mb - e)) & ((1u << mb) - 1u);
      bucket = 1u + ((uint32_t)e << mb) + m;
    }
    const int kb = sort_key_bits(sort_mode) - 1;
    sort_key[g] = (single << kb) | (((1u << kb) - 1u) - bucket);
  }
  sort_idx[g] = (int32_t)g;
}

__device__ __forceinline__ int adc_f(const crv_params &p, float v) {
  float x = v * (float)p.adc_conversion + (float)p.pedestal + 0.5f;
  x = fminf(fmaxf(x, -1.f), (float)p.max_adc + 1.f);
  const float a = truncf(x);
  return (int)fminf(fmaxf(a, (float)p.min_adc), (float)p.max_adc);
}

// ---- pulse fit: Gumbel A exp(-z - e^-z), z = (t - mu) / beta, in sine-mapped coordinates ----
struct FSum {
  float chi2, jtj[6], jtr[3];
};

__device__ __forceinline__ void fparams(const float lo[3], const float hi[3], const float u[3], float par[3],
                                        float dpdu[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float span = hi[i] - lo[i];
    float sn, cs;
    __sincosf(u[i], &sn, &cs);
    par[i] = lo[i] + span * (sn + 1.f) * 0.5f;
    dpdu[i] = span * 0.5f * cs;
  }
}

// chi2, J^T J and J^T r of the Gumbel model at u over samples fs..fe; sample q is y[(q - q0) * ys]
__device__ __forceinline__ void feval(const float *y, int64_t ys, int q0, int fs, int fe, const float lo[3],
                                      const float hi[3], const float u[3], FSum &s) {
  float par[3], dpdu[3];
  fparams(lo, hi, u, par, dpdu);
  const float A = par[0], mu = par[1], ib = 1.f / par[2];
  s.chi2 = 0.f;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.jtj[i] = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) s.jtr[i] = 0.f;
  for (int q = fs; q <= fe; ++q) {
    const float z = ((float)q * PERIODF - mu) * ib;
    const float ez = __expf(-z);
    const float g = __expf(-z - ez);
    const float fv = A * g;
    const float r = y[(q - q0) * ys] - fv;
    const float dfdz = fv * (ez - 1.f);
    const float J0 = g * dpdu[0], J1 = -dfdz * ib * dpdu[1], J2 = -dfdz * z * ib * dpdu[2];
    s.chi2 += r * r;
    s.jtr[0] += J0 * r;
    s.jtr[1] += J1 * r;
    s.jtr[2] += J2 * r;
    s.jtj[0] += J0 * J0;
    s.jtj[1] += J0 * J1;
    s.jtj[2] += J0 * J2;
    s.jtj[3] += J1 * J1;
    s.jtj[4] += J1 * J2;
    s.jtj[5] += J2 * J2;
  }
}

// (S + diag(add)) x = b in FP64 for the symmetric 3 x 3 S: cofactors and one reciprocal of the determinant (the FP64
// chain eliminates with partial pivoting, as the reference's solve does; here the inputs carry FP32 rounding, far
// above what the two methods' FP64 differences reach); false on a zero determinant
__device__ __forceinline__ bool fsolve(const float jtj[6], const double add[3], const float bv[3], double x[3]) {
  const double a00 = (double)jtj[0] + add[0], a01 = jtj[1], a02 = jtj[2], a11 = (double)jtj[3] + add[1];
  const double a12 = jtj[4], a22 = (double)jtj[5] + add[2];
  const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
  const double det = a00 * c00 + a01 * c01 + a02 * c02;
  if (det == 0.0) return false;
  const double c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
  const double id = 1.0 / det, b0 = bv[0], b1 = bv[1], b2 = bv[2];
  x[0] = (c00 * b0 + c01 * b1 + c02 * b2) * id;
  x[1] = (c01 * b0 + c11 * b1 + c12 * b2) * id;
  x[2] = (c02 * b0 + c12 * b1 + c22 * b2) * id;
  return true;
}

// extreme eigenvalues of the symmetric 3 x 3 normal matrix: six cyclic Jacobi sweeps
__device__ __forceinline__ void feig(const float jtj[6], float &emin, float &emax) {
  float a[3][3] = {{jtj[0], jtj[1], jtj[2]}, {jtj[1], jtj[3], jtj[4]}, {jtj[2], jtj[4], jtj[5]}};
  const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 6; ++sweep) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = P[k], q = Q[k], r = 3 - p - q;
      const float apq = a[p][q];
      if (apq == 0.f) continue;
      const float th = (a[q][q] - a[p][p]) / (2.f * apq);
      const float tt = (th >= 0.f ? 1.f : -1.f) / (fabsf(th) + sqrtf(th * th + 1.f));
      const float c = 1.f / sqrtf(tt * tt + 1.f), s = tt * c;
      const float arp = a[r][p], arq = a[r][q];
      a[p][p] -= tt * apq;
      a[q][q] += tt * apq;
      a[p][q] = a[q][p] = 0.f;
      a[r][p] = a[p][r] = c * arp - s * arq;
      a[r][q] = a[q][r] = s * arp + c * arq;
    }
  }
  emin = fminf(fminf(a[0][0], a[1][1]), a[2][2]);
  emax = fmaxf(fmaxf(a[0][0], a[1][1]), a[2][2]);
}


// a pulse slot of a pool run: the fit's results, read back whole by the pulse-table kernel
struct alignas(16) Slot {
  double tp, mu;  // pulse time without the DB offset, fitted Gumbel location (absolute, ns)
  float pe, height, beta, chi2, ped, fa, fb;
  uint32_t sf;  // seq | failed << 16 | zerondf << 17 | pedfromdb << 18
};

struct ClDesc {  // one charge cluster of a SiPM whose waveform is kept: its log entries [q0, q1)
  int64_t lb;     // the SiPM's log base (goff of its group + its lane)
  int32_t g, q0, q1, pad_;
};

struct Out {
  // fired-pixel log: entries (time - tbase of the SiPM, overvoltage) in chunks of two (16 bytes); per group of 64
  // SiPMs in queue order, chunk c of the group's lane l at goff + 64 c + l (the lanes of a wavefront write and read
  // neighbouring chunks); a group holds gcap entries (even) per lane
  const int64_t *goff, *gcap;
  int64_t log0;     // goff of the pass's first group: the pass's log starts there
  float4 *flog;
  double *tbase;     // per SiPM: the time of its first step
  float *lastt;      // per persistent lane: time offset of each pixel's last firing, [pixel][grid lane]
  // photon streams of the (step, SiPM) pairs of counters with several steps
  float *gS;
  double *gHead, *gTb;
  uint64_t *gRJ;  // per stream: step ordinal in its event << 32 | photon rank (ranks stay below 65536)
  // cascade lists per persistent lane [slot][grid lane]
  uint16_t *xcur, *xcand;
  uint8_t *xmult;
  int64_t xstride;
  // record pool: window_samples ADC words, label and FEB tick of a cluster with digis (label -1: unused slot)
  uint16_t *rec_adc;
  int64_t *rec_label;
  int32_t *rec_tick;
  int32_t *rec_count;
  int32_t *rec_used;  // records that hold a waveform with digis (the others stay reserved with label -1)
  int64_t rec_cap;
  // run pool: the reco waveforms (record, first sample, samples, run number (in its cluster, then in the SiPM after
  // kb_run_rank), SiPM, start TDC)
  int32_t *run_rec, *run_i0, *run_n, *run_k;
  int64_t *run_g, *run_tdc;
  int32_t *run_count;
  int64_t run_cap;
  int32_t *nrun;  // per SiPM
  // charge clusters whose waveform is kept: per SiPM its first descriptor and their count; the descriptor pool (the
  // work units of kb_wrec) and the descriptors reserved so far; per descriptor its reco waveforms and their offset
  // within the SiPM; per run its descriptor
  int32_t *cd0, *ncls;
  struct ClDesc *cdesc;
  int32_t *cdesc_count;
  int64_t cdesc_cap;
  int32_t *ncr, *crl;
  int32_t *run_c;
  // pulse slots per pool run (one 48-byte record each), and per final run the pulse count and the pool run
  int pcap;
  Slot *slot;
  int32_t *npr, *rmap;
  float *yscr;  // fit samples [sample][pool run]
};

// work queue of a persistent grid: a workgroup takes `tile` SiPMs (in `order`) per grab
struct Queue {
  int *ctr;
  int tile;
  int64_t n;  // one past the last queue position
  const int32_t *order;
  int64_t start;  // first queue position (a multiple of 64)
  const int32_t *tdev = nullptr;  // fixed-capacity calls: the hit counters, n limited to 4 of them per counter
  const int *skip = nullptr;      // fixed-capacity calls: nonzero when the call exceeded a capacity
};

// the queue of this launch: fixed-capacity calls read their extent from device memory; false: nothing to do
__device__ __forceinline__ bool resolve(Queue &q) {
  if (q.skip && *q.skip) return false;
  if (q.tdev) q.n = min(q.n, 4 * (int64_t)(*q.tdev));
  return true;
}

__device__ __forceinline__ bool grab(const Queue &qu, int *tile_base, int64_t &base) {
  if (threadIdx.x == 0) *tile_base = atomicAdd(qu.ctr, qu.tile);
  __syncthreads();
  base = qu.start + *tile_base;
  __syncthreads();
  return base < qu.n;
}

__device__ __forceinline__ bool bget(const uint32_t *bm, int lane, int pix) {
  return (bm[(pix >> 5) * 64 + lane] >> (pix & 31)) & 1u;
}
__device__ __forceinline__ void bset(uint32_t *bm, int lane, int pix) { bm[(pix >> 5) * 64 + lane] |= 1u << (pix & 31); }

// a pixel whose bit is set in a bitmap that started as the dead mask: dead (lost) or fired before in this SiPM
__device__ __forceinline__ bool dead_px(const Tables &t, int pix) {
  if (t.dead_w0 < 0) return (t.dead[pix >> 5] >> (pix & 31)) & 1u;
  const int k = (pix >> 5) - t.dead_w0;  // the mask's nonzero words are in registers: no load
  uint32_t w = 0u;
#pragma unroll
  for (int i = 0; i < 8; ++i) w = k == i ? t.dead_win[i] : w;
  return (w >> (pix & 31)) & 1u;
}

// heads visited in increasing stream index: (h1, i1) the earliest (first index on ties), (h2, i2) the earliest of the
// others
__device__ __forceinline__ void earliest2(double h, int k, double &h1, int &i1, double &h2, int &i2) {
  const bool lt1 = h < h1, lt2 = h < h2;
  h2 = lt1 ? h1 : lt2 ? h : h2;
  i2 = lt1 ? i1 : lt2 ? k : i2;
  h1 = lt1 ? h : h1;
  i1 = lt1 ? k : i1;
}

// append a firing to the SiPM's log and to the per-pixel last-fire times
// (an even entry waits in registers, pt / pv, and goes out with the next one as one 16-byte chunk)
__device__ __forceinline__ void log_fire(const Out &o, int64_t lb, int64_t &nlog, float &pt, float &pv, float *lastt,
                                         int64_t xs, int pix, float toff, float v) {
  if (nlog & 1) {
    o.flog[lb + (nlog >> 1) * 64] = make_float4(pt, pv, toff, v);
  } else {
    pt = toff;
    pv = v;
  }
  lastt[pix * xs] = toff;
  ++nlog;
}

// the kept charge clusters of one SiPM's fired-pixel log (entries [0, nf) at lb): clusters split at gaps of more than
// `gap`, kept when nq * vmax >= vmin (nq: the cluster's charges); the first CL_REG of them in (qs, qe), the ones from
// `emit_from` on written to dst; returns the number kept
constexpr int CL_REG = 4;
__device__ __forceinline__ int scan_clusters(const Out &o, int64_t lb, int nf, double tb0, double gap, float vmax,
                                             float vmin, int32_t qs[CL_REG], int32_t qe[CL_REG], int emit_from,
                                             ClDesc *dst, int32_t g) {
  int q = 0, q_open = 0, nq = 0, m = 0;
  bool open = false;
  double t_last = 0.0;
  float4 ck = make_float4(0.f, 0.f, 0.f, 0.f);
  if (nf > 0) ck = o.flog[lb];
  auto emit = [&](int q1) {
    if ((float)nq * vmax >= vmin) {
#pragma unroll
      for (int k = 0; k < CL_REG; ++k)
        if (k == m) {
          qs[k] = q_open;
          qe[k] = q1;
        }
      if (dst && m >= emit_from) dst[m] = ClDesc{lb, g, q_open, q1, 0};
      ++m;
    }
  };
  while (q < nf) {
    const float tqf = (q & 1) ? ck.z : ck.x;
    const double tq = tb0 + (double)tqf;
    int n = 1;
    for (;;) {
      const int e = q + n;
      if (e >= nf) break;
      if ((e & 1) == 0) ck = o.flog[lb + (int64_t)(e >> 1) * 64];
      if (((e & 1) ? ck.z : ck.x) != tqf) break;
      ++n;
    }
    if (open && tq - t_last > gap) {
      emit(q);
      open = false;
    }
    if (!open) {
      open = true;
      q_open = q;
      nq = 0;
    }
    nq += n;
    t_last = tq;
    q += n;
  }
  if (open) emit(nf);
  return m;
}

// kb_sipm, one lane per SiPM: photons of the counter's steps in time order -> pixel, recovery, firing, cross-talk
// cascade -> the SiPM's fired-pixel log (time, overvoltage, pixel), in firing order, and a descriptor for each of its
// kept charge clusters (cd0 / ncls: the SiPM's first descriptor and their count).
__global__ __launch_bounds__(64) void kb_sipm(crv_params p, Tables t, Derived dv, BStep st,
                                              const uint64_t *__restrict__ task_key,
                                              const int32_t *__restrict__ task_off, const int32_t *__restrict__ perm,
                                              const int32_t *__restrict__ evfirst, const int32_t *__restrict__ nph,
                                              Out o, Queue qu, int32_t *__restrict__ nfired, int *__restrict__ flags) {
  __shared__ uint32_t bm[CRV_DEAD_WORDS * 64];
  __shared__ int tile_base;
  if (!resolve(qu)) return;
  const int lane = threadIdx.x;
  const float Vov = (float)p.overvoltage, p1 = (float)p.p1;
  const float inv_tau = 1.f / (float)p.time_constant, inv_p2 = 1.f / (float)p.p2;
  const float pf_full = p1 * (1.f - __expf(-Vov * inv_p2));
  const float qct = (float)dv.q_ct, tauf = (float)p.tau;
  const DrawR dr{(uint32_t)p.seed, p.evkey_hi};
  const float inv_ny = 1.f / (float)t.map_ny;
  const int64_t xb = (int64_t)blockIdx.x * 64 + lane;
  uint16_t *xcur = o.xcur + xb, *xcand = o.xcand + xb;
  uint8_t *xmult = o.xmult + xb;
  const int64_t xs = o.xstride;
  float *lastt = o.lastt + xb;  // read only where the pixel's bitmap bit is set: no clearing between SiPMs
  // charge clusters: the split rule of the waveform window (a gap of more than single_pe_max_time + 4 periods after
  // the last logged time closes a cluster), kept when its charges reach the zero-suppression threshold
  const double gap = p.single_pe_max_time + 4.0 * PERIODD;
  const float vmax = (float)t.vmax, vmin = (float)p.min_voltage;
  int64_t base = 0;
  while (grab(qu, &tile_base, base)) {
    for (int64_t gq0 = base; gq0 < base + qu.tile && gq0 < qu.n; gq0 += 64) {
      const int64_t gq = gq0 + lane;
      const int64_t g = gq < qu.n ? (qu.order ? (int64_t)qu.order[gq] : gq) : -1;
      const int64_t lb = o.goff[gq0 >> 6] - o.log0 + lane;
      double tbase = 0.0;
      int64_t nlog = 0;
      int m = 0;  // kept clusters; the first CL_REG of them in (qs, qe)
      int32_t qs[CL_REG], qe[CL_REG];
      do {  // the lane's SiPM (none past the end of the queue)
        if (g < 0) break;
        const int64_t tk = g >> 2;
        const int s = (int)(g & 3);
        const uint64_t key = task_key[tk];
        const uint32_t ev = (uint32_t)(key >> 13);
        const int ch = (int)(key & 8191u) * 4 + s;
        if (ch >= t.n_channels || (t.status[ch] & 0b101) != 0) {
          nfired[g] = 0;
          break;
        }
        const int a = task_off[tk], b = task_off[tk + 1];
        const int nstr = b - a;
        // the bitmap starts as the dead-pixel mask: a clear bit (the usual case) is a live pixel that has not fired,
        // so only a set bit needs the dead table (dead: the photon is lost) or the pixel's last firing
        for (int w = 0; w < CRV_DEAD_WORDS; ++w) bm[w * 64 + lane] = t.dead[w];
        const int64_t lcap = o.gcap[gq0 >> 6];
        tbase = st.t[perm[a]];  // log times are FP32 offsets from it (they span microseconds)
        o.tbase[g] = tbase;
        int q_open = 0;      // first entry of the open cluster
        double c_last = 0.0; // time of the last logged entry
        auto close_cluster = [&](int q1) {
          if ((float)(q1 - q_open) * vmax >= vmin) {
  #pragma unroll
            for (int k = 0; k < CL_REG; ++k)
              if (k == m) {
                qs[k] = q_open;
                qe[k] = q1;
              }
            ++m;
          }
        };
        float pend_t = 0.f, pend_v = 0.f;  // the log's pending even entry
        // photon streams: the current stream in registers; with several steps the other streams in global memory
        // (gS, gTb, gHead, gRJ) and the earliest of their heads (th, from stream kth) in registers. Photons come from
        // the current stream while its head is the earliest (on a tie, from the lower step); otherwise its state goes
        // back to memory and a scan of the heads makes the earliest the current stream. Streams of separate steps
        // seldom overlap in time, so a SiPM scans about once per step instead of once per photon.
        float S1 = 0.f;
        double head1 = 0.0, tb1 = 0.0, th = INFINITY;
        int r1 = 0, N1 = 0, cur = -1, kth = 0;
        uint32_t j1 = 0;
        U4 dq1{0u, 0u, 0u, 0u};  // the delay draw holding the current stream's next four exponentials
        for (int k = a; k < b; ++k) {
          const int64_t gi = (int64_t)k * 4 + s;
          const int N = nph[gi];
          if (nstr > 1 && N <= 0) {
            o.gRJ[gi] = 0;
            o.gHead[gi] = INFINITY;  // an empty or exhausted stream's head is +inf
          }
          if (N <= 0) continue;
          const int32_t i = perm[k];
          const uint32_t j = (uint32_t)(i - evfirst[i]);
          const double z = st.z[i];
          const double d = (s & 1) == 0 ? z + p.half_length : p.half_length - z;
          const double tb = st.t[i] + d * p.inv_speed;
          const U4 d0 = dr(KIND_DELAY, ev, (uint32_t)ch, j, 0u);
          const float S = -__logf(u32f(d0.x)) / (float)N;
          if (nstr == 1) {
            cur = k;
            dq1 = d0;
            N1 = N;
            j1 = j;
            tb1 = tb;
            S1 = S;
            head1 = tb + (double)(tauf * S);
          } else {
            o.gS[gi] = S;
            o.gTb[gi] = tb;
            o.gHead[gi] = tb + (double)(tauf * S);
            o.gRJ[gi] = (uint64_t)j << 32;
          }
        }
        bool bad = false;
        for (;;) {
          // next photon: the earliest stream head (lowest step first on ties)
          bool scan = false;
          if (nstr == 1) {
            if (r1 >= N1) break;  // its one stream is exhausted
          } else {
            scan = !(r1 < N1 && (head1 < th || (head1 == th && cur < kth)));
          }
          if (scan) {
            if (cur >= 0) {  // the current stream back to memory
              const int64_t gi = (int64_t)cur * 4 + s;
              o.gS[gi] = S1;
              o.gHead[gi] = r1 < N1 ? head1 : INFINITY;
              o.gRJ[gi] = (uint64_t)j1 << 32 | (uint32_t)r1;
            }
            // the earliest head (first index on ties) and the earliest of the others; one load per stream, eight in
            // flight at a time
            double h1 = INFINITY, h2 = INFINITY;
            int i1 = -1, i2 = 0;
            int k = a;
            for (; k + 8 <= b; k += 8) {
              double h[8];
  #pragma unroll
              for (int q = 0; q < 8; ++q) h[q] = o.gHead[(int64_t)(k + q) * 4 + s];
  #pragma unroll
              for (int q = 0; q < 8; ++q) earliest2(h[q], k + q, h1, i1, h2, i2);
            }
            for (; k < b; ++k) earliest2(o.gHead[(int64_t)k * 4 + s], k, h1, i1, h2, i2);
            if (i1 < 0) break;  // every stream is exhausted
            th = h2;
            kth = i2;
            cur = i1;
            const int64_t gi = (int64_t)cur * 4 + s;
            const uint64_t rj = o.gRJ[gi];  // one load for the step ordinal and the rank
            j1 = (uint32_t)(rj >> 32);
            r1 = (int)(uint32_t)rj;
            N1 = nph[gi];
            S1 = o.gS[gi];
            tb1 = o.gTb[gi];
            head1 = h1;
            dq1 = dr(KIND_DELAY, ev, (uint32_t)ch, j1, (uint32_t)(r1 >> 2));
          }
          const double tph = head1;
          if (tph > p.photon_window_hi) break;  // later photons of every stream are later still
          const float toff = (float)(tph - tbase);
          // advance the stream
          const uint32_t j = j1;
          const int r = r1;
          if (r1 + 1 < N1) {
            const int rn = r1 + 1;
            if ((rn & 3) == 0) dq1 = dr(KIND_DELAY, ev, (uint32_t)ch, j1, (uint32_t)(rn >> 2));
            const uint32_t x = (rn & 3) == 0 ? dq1.x : (rn & 3) == 1 ? dq1.y : (rn & 3) == 2 ? dq1.z : dq1.w;
            S1 += (-__logf(u32f(x))) / (float)(N1 - rn);
            head1 = tb1 + (double)(tauf * S1);
          }
          r1 += 1;
          // pixel of the photon (photon-map CDF), firing with the recovered overvoltage
          const U4 u = dr(KIND_PIXEL, ev, (uint32_t)ch, j, pack_photon((uint32_t)r, 0u, 0u));
          const double xc = unif(u.x) * t.cdf_total;
          int bin;
          if (t.map_uniform) {
            bin = min((int)xc, t.ncdf - 1);
          } else {
            int lo = 0, hi = t.ncdf;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (t.cdf[mid] <= xc) lo = mid + 1;
              else hi = mid;
            }
            bin = min(lo, t.ncdf - 1);
          }
          int px = (int)(((float)bin + 0.5f) * inv_ny);  // bin < 1600: the float quotient's floor is exact
          int py = bin - px * t.map_ny;
          if (p.pixel_lrint) {
            px += u32f(u.y) >= 0.5f ? 1 : 0;
            py += u32f(u.z) >= 0.5f ? 1 : 0;
          }
          const int pix = px * CRV_NPIX + py;
          const bool hit_set = bget(bm, lane, pix);
          if (hit_set && dead_px(t, pix)) continue;
          float v = Vov;
          if (hit_set) v = Vov * (1.f - __expf(-(toff - lastt[pix * xs]) * inv_tau));
          const float pf = v == Vov ? pf_full : p1 * (1.f - __expf(-v * inv_p2));
          if (!(u32f(u.w) < pf)) continue;
          if (nlog >= lcap) {
            bad = true;
            break;
          }
          {  // the photon's firings (its pixel and cross-talk) share its time
            const double tq = tbase + (double)toff;
            if (nlog > 0 && tq - c_last > gap) {
              close_cluster((int)nlog);
              q_open = (int)nlog;
            }
            c_last = tq;
          }
          log_fire(o, lb, nlog, pend_t, pend_v, lastt, xs, pix, toff, v);
          bset(bm, lane, pix);
          // cross-talk cascade, breadth-first, candidates of a generation in ascending pixel order. While a
          // generation's pixels stay inside the 7 x 7 window around the photon's pixel, the generation, its candidates
          // and their multiplicities are bit masks over the window in registers (bit (dx + 3) * 7 + dy + 3; ascending
          // bit = ascending pixel); a generation with a pixel on the window's border continues in the per-lane lists
          // in global scratch below. Both visit the same draws in the same order.
          const int cx = pix / CRV_NPIX, cy = pix % CRV_NPIX;
          uint64_t cur = 1ull << 24;  // generation 0: the photon's pixel at the window's centre
          int gnum = 0;
          bool window = true;
          while (!bad) {
            if (cur & WIN_BORDER) {
              window = false;  // neighbours could leave the window
              break;
            }
            uint64_t m0 = 0, m1 = 0, m2 = 0;  // candidate multiplicity, bit planes of a 3-bit count per position
            for (uint64_t rem = cur; rem; rem &= rem - 1) {
              const int bq = __ffsll((long long)rem) - 1;
              const int fx = cx + bq / 7 - 3, fy = cy + bq % 7 - 3;
              const U4 ux = dr(KIND_XTALK, ev, (uint32_t)ch, j,
                               pack_photon((uint32_t)r, (uint32_t)gnum, (uint32_t)(fx * CRV_NPIX + fy)));
              uint64_t add = 0;
              if (fx > 0 && u32f(ux.x) < qct) add |= 1ull << (bq - 7);
              if (fx + 1 < p.n_pixels_x && u32f(ux.y) < qct) add |= 1ull << (bq + 7);
              if (fy > 0 && u32f(ux.z) < qct) add |= 1ull << (bq - 1);
              if (fy + 1 < p.n_pixels_y && u32f(ux.w) < qct) add |= 1ull << (bq + 1);
              const uint64_t c0 = m0 & add;
              m0 ^= add;
              const uint64_t c1 = m1 & c0;
              m1 ^= c0;
              m2 ^= c1;
            }
            const uint64_t cand = m0 | m1 | m2;
            if (!cand) break;
            if (__popcll(cand) > CMAXB || gnum + 1 > MAXGEN) {
              atomicAdd(&flags[F_CASCADE], 1);
              bad = true;
              break;
            }
            ++gnum;
            uint64_t nxt = 0;
            for (uint64_t rem = cand; rem; rem &= rem - 1) {
              const int bq = __ffsll((long long)rem) - 1;
              const int c = (cx + bq / 7 - 3) * CRV_NPIX + cy + bq % 7 - 3;
              const int m = (int)((m0 >> bq) & 1u) + 2 * (int)((m1 >> bq) & 1u) + 4 * (int)((m2 >> bq) & 1u);
              const bool c_set = bget(bm, lane, c);
              if (c_set && dead_px(t, c)) continue;
              float vc = Vov;
              if (c_set) vc = Vov * (1.f - __expf(-(toff - lastt[c * xs]) * inv_tau));
              const float q1 = vc == Vov ? 1.f - pf_full : 1.f - p1 * (1.f - __expf(-vc * inv_p2));
              float qm = q1;
              for (int mm = 2; mm <= m; ++mm) qm *= q1;
              const float uc =
                  u32f(dr(KIND_XTFIRE, ev, (uint32_t)ch, j, pack_photon((uint32_t)r, (uint32_t)gnum, (uint32_t)c)).x);
              if (!(uc < 1.f - qm)) continue;
              if (nlog >= lcap) {
                bad = true;
                break;
              }
              log_fire(o, lb, nlog, pend_t, pend_v, lastt, xs, c, toff, vc);
              bset(bm, lane, c);
              nxt |= 1ull << bq;
            }
            cur = nxt;
            if (!cur) break;
          }
          if (bad) break;
          if (window) continue;
          int ncur = 0;
          for (uint64_t rem = cur; rem; rem &= rem - 1, ++ncur) {
            const int bq = __ffsll((long long)rem) - 1;
            xcur[ncur * xs] = (uint16_t)((cx + bq / 7 - 3) * CRV_NPIX + cy + bq % 7 - 3);
          }
          while (ncur > 0 && !bad) {
            int ncand = 0;
            for (int ai = 0; ai < ncur; ++ai) {
              const int f = (int)xcur[ai * xs];
              const int fx = f / CRV_NPIX, fy = f % CRV_NPIX;
              const U4 ux = dr(KIND_XTALK, ev, (uint32_t)ch, j, pack_photon((uint32_t)r, (uint32_t)gnum, (uint32_t)f));
              const uint32_t uw[4] = {ux.x, ux.y, ux.z, ux.w};
  #pragma unroll
              for (int nb = 0; nb < 4; ++nb) {
                const int nx = nb == 0 ? fx - 1 : nb == 1 ? fx + 1 : fx;
                const int ny = nb == 2 ? fy - 1 : nb == 3 ? fy + 1 : fy;
                const bool ok =
                    nb == 0 ? fx > 0 : nb == 1 ? fx + 1 < p.n_pixels_x : nb == 2 ? fy > 0 : fy + 1 < p.n_pixels_y;
                if (!ok || !(u32f(uw[nb]) < qct)) continue;
                const int c = nx * CRV_NPIX + ny;
                int at = 0;
                while (at < ncand && xcand[at * xs] < c) ++at;
                if (at < ncand && xcand[at * xs] == c) {
                  xmult[at * xs] += 1;
                } else if (ncand < CMAXB) {
                  for (int q = ncand; q > at; --q) {
                    xcand[q * xs] = xcand[(q - 1) * xs];
                    xmult[q * xs] = xmult[(q - 1) * xs];
                  }
                  xcand[at * xs] = (uint16_t)c;
                  xmult[at * xs] = 1;
                  ++ncand;
                } else {
                  atomicAdd(&flags[F_CASCADE], 1);
                  bad = true;
                }
              }
            }
            if (ncand == 0 || bad) break;
            if (++gnum > MAXGEN) {
              atomicAdd(&flags[F_CASCADE], 1);
              bad = true;
              break;
            }
            ncur = 0;
            for (int ai = 0; ai < ncand; ++ai) {
              const int c = xcand[ai * xs];
              const int m = xmult[ai * xs];
              const bool c_set = bget(bm, lane, c);
              if (c_set && dead_px(t, c)) continue;
              float vc = Vov;
              if (c_set) vc = Vov * (1.f - __expf(-(toff - lastt[c * xs]) * inv_tau));
              const float q1 = vc == Vov ? 1.f - pf_full : 1.f - p1 * (1.f - __expf(-vc * inv_p2));
              float qm = q1;
              for (int mm = 2; mm <= m; ++mm) qm *= q1;
              const float uc =
                  u32f(dr(KIND_XTFIRE, ev, (uint32_t)ch, j, pack_photon((uint32_t)r, (uint32_t)gnum, (uint32_t)c)).x);
              if (!(uc < 1.f - qm)) continue;
              if (nlog >= lcap) {
                bad = true;
                break;
              }
              log_fire(o, lb, nlog, pend_t, pend_v, lastt, xs, c, toff, vc);
              bset(bm, lane, c);
              xcur[ncur * xs] = (uint16_t)c;
              ++ncur;
            }
          }
          if (bad) break;
        }
        if (nlog & 1) o.flog[lb + (nlog >> 1) * 64] = make_float4(pend_t, pend_v, 0.f, 0.f);
        if (bad) {
          atomicAdd(&flags[F_OVERFLOW], 1);  // the batch is run again with larger logs
          m = 0;
        } else if (nlog > 0) {
          close_cluster((int)nlog);
        }
        nfired[g] = (int32_t)nlog;
      } while (false);
      // the descriptors of the wavefront's 64 SiPMs, reserved together in lane order (kb_wrec's lanes then read
      // neighbouring log chunks)
      int incl = m;  // inclusive prefix over the lanes
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
      }
      const int tot = __shfl(incl, 63);
      int wb = 0;
      if (lane == 0 && tot) wb = atomicAdd(o.cdesc_count, tot);
      wb = __shfl(wb, 0);
      const int32_t b0 = wb + incl - m;
      if (g >= 0) {
        if ((int64_t)b0 + m <= o.cdesc_cap) {  // else the pool is full: the call runs again with a larger one
#pragma unroll
          for (int k = 0; k < CL_REG; ++k)
            if (k < m) o.cdesc[b0 + k] = ClDesc{lb, (int32_t)g, qs[k], qe[k], 0};
          if (m > CL_REG)  // the rest from the log just written
            scan_clusters(o, lb, (int)nlog, tbase, gap, vmax, vmin, qs, qe, CL_REG, o.cdesc + b0, (int32_t)g);
        }
        o.cd0[g] = b0;
        o.ncls[g] = m;
      }
    }
  }
}


// ---- waveform machine (kb_wrec) ----
struct Wave {
  float win[RINGB];        // open samples E .. E + RINGB - 1
  U4 nz;                   // noise draw of the sample pair holding E - 1
  double t_start, t_first, t_last;
  int32_t kf, nq, used, E, next_i, nd;
  uint32_t above;          // bit k: sample E - 1 - k above threshold
  uint32_t dmask[DWORDS];  // bit i: a digi starts at sample i
  int32_t rec;             // reserved record slot (-1: none)
  bool open;
  int32_t overlong;
  uint32_t ab[4];          // ADC words of the record's current 8-sample chunk
};

struct Lane {
  uint32_t ev;
  int ch;
  int64_t evkey;
  double td;  // TDC0 of the SiPM's FEB
  int64_t task;
  int nrun;   // reco waveforms of the cluster so far
  int32_t cl; // the cluster's descriptor (kb_wrec)
  int32_t used;  // records with digis the lane wrote (added to rec_used once, when the lane is done)
  int W;
  const float *tmpl;
};

// ADC word a of sample J of the lane's record, gathered in registers and written eight samples (16 bytes) at a time;
// a record's samples come in order from 0 (records whose length is not a multiple of 8: word by word)
__device__ __forceinline__ void put_adc(const Out &o, const Lane &L, Wave &w, int J, uint32_t a) {
  if ((L.W & 7) != 0) {
    o.rec_adc[(int64_t)w.rec * L.W + J] = (uint16_t)a;
    return;
  }
  const int k = (J & 7) >> 1;
  const uint32_t bits = (a & 0xFFFFu) << ((J & 1) * 16);
#pragma unroll
  for (int q = 0; q < 4; ++q) w.ab[q] |= q == k ? bits : 0u;
  if ((J & 7) == 7) {
    *(uint4 *)(o.rec_adc + (int64_t)w.rec * L.W + (J & ~7)) = make_uint4(w.ab[0], w.ab[1], w.ab[2], w.ab[3]);
    w.ab[0] = w.ab[1] = w.ab[2] = w.ab[3] = 0u;
  }
}

// finalise sample E: noise, ADC word, zero-suppression decision for sample E - 2
__device__ __forceinline__ void wave_step(const crv_params &p, const Lane &L, const Out &o, Wave &w, bool at_close,
                                          int length) {
  const int J = w.E++;
  float v = w.win[0];
#pragma unroll
  for (int k = 0; k < RINGB - 1; ++k) w.win[k] = w.win[k + 1];
  w.win[RINGB - 1] = 0.f;
  const bool valid = !at_close || J < length;
  if (valid && p.noise_on) {
    if ((J & 1) == 0)
      w.nz = DrawR{(uint32_t)p.seed, p.evkey_hi}(KIND_NOISE, L.ev, (uint32_t)L.ch, (uint32_t)w.kf, (uint32_t)(J >> 1));
    const float u1 = (J & 1) ? u32f(w.nz.z) : u32f(w.nz.x), u2 = (J & 1) ? u32f(w.nz.w) : u32f(w.nz.y);
    v += (float)p.noise * sqrtf(-2.f * __logf(u1)) * __cosf(6.28318531f * u2);
  }
  if (J < L.W && w.rec >= 0) put_adc(o, L, w, J, (uint32_t)adc_f(p, valid ? v : 0.f));
  const bool ab = valid && v > (float)p.min_voltage;
  w.above = (w.above << 1) | (ab ? 1u : 0u);
  const int i = J - 2;  // zero suppression of sample i needs samples i - 2 .. i + 2
  const bool i_valid = i >= 0 && (!at_close || i < length);
  if (i_valid && (w.above & 0x1Fu) != 0u && w.next_i <= i) {
    const double tstart = w.t_start + (double)i * PERIODD;
    if (tstart >= p.window_lo && tstart <= p.window_hi) {  // i < length <= W - samples_zs < 32 * DWORDS
      const uint32_t bit = 1u << (i & 31);
#pragma unroll
      for (int q = 0; q < DWORDS; ++q) w.dmask[q] |= q == (i >> 5) ? bit : 0u;  // a select per word: stays in VGPRs
      ++w.nd;
      w.next_i = i + p.samples_zs;
    }
  }
}

__device__ __forceinline__ void wave_open(const Lane &L, const Out &o, Wave &w, double t) {
  w.open = true;
  w.t_first = t;
  w.t_last = t;
  w.kf = (int32_t)ceil((t - PERIODD - L.td) / PERIODD);
  w.t_start = (double)w.kf * PERIODD + L.td;
  w.nq = 0;
  w.used = 0;
  w.E = 0;
  w.ab[0] = w.ab[1] = w.ab[2] = w.ab[3] = 0u;
  w.above = 0u;
  w.next_i = 0;
  w.nd = 0;
#pragma unroll
  for (int k = 0; k < RINGB; ++k) w.win[k] = 0.f;
#pragma unroll
  for (int k = 0; k < DWORDS; ++k) w.dmask[k] = 0u;
  if (w.rec < 0) {  // a record slot; a cluster without digis leaves it to the next cluster
    const int r = atomicAdd(o.rec_count, 1);
    w.rec = r < o.rec_cap ? r : -1;
  }
}

// nf charges at time tq with summed overvoltage vsum (a fired pixel and its cross-talk): template sums
__device__ __forceinline__ void wave_add(const crv_params &p, const Tables &t, const Lane &L, const Out &o, Wave &w,
                                         double tq, float vsum, int nf, int nreach) {
  const int lmax = L.W - p.samples_zs;
  w.nq += nf;
  w.t_last = tq;
  const float dt = (float)(tq - w.t_start);
  const int j0 = (int)fmaxf(ceilf(dt / PERIODF), 0.f);
  while (w.E < j0 && w.E < lmax) wave_step(p, L, o, w, false, 0);
  if (j0 >= lmax) {  // every sample the charges reach is past the window
    w.overlong = 1;
    return;
  }
  // here E == j0: j0 never decreases within a cluster and E only advances to it. Sample j0 + r sees the template
  // at t = stt0 + r * 12.5 (stt0 in [0, 12.5)), index rint(t / precision), while that index is inside the template.
  const float q = (float)p.capacitance * vsum / (float)p.single_pe_reference_charge;
  const float inv_prec = 1.f / (float)p.single_pe_precision;
  const float b0 = ((float)j0 * PERIODF - dt) * inv_prec, db = PERIODF * inv_prec;
  const float nt = (float)t.ntmpl;
  int rr = 0;  // samples reached
#pragma unroll
  for (int r = 0; r < RINGB; ++r) {
    const float ti = rintf(fmaf((float)r, db, b0));
    if (r < nreach && ti < nt) {
      w.win[r] += L.tmpl[(int)ti] * q;
      rr = r + 1;
    }
  }
  if (j0 + rr > lmax) {  // the part past the window is dropped
    w.overlong = 1;
#pragma unroll
    for (int r = 0; r < RINGB; ++r)
      if (j0 + r >= lmax) w.win[r] = 0.f;
    rr = lmax - j0;
  }
  w.used = max(w.used, j0 + rr);
}

__device__ __forceinline__ uint32_t dword_at(const uint32_t m[DWORDS], int q) {
  uint32_t v = 0u;
#pragma unroll
  for (int k = 0; k < DWORDS; ++k) v = q == k ? m[k] : v;
  return v;
}

// first set bit >= pos of the digi mask, or -1
__device__ __forceinline__ int next_digi(const uint32_t m[DWORDS], int pos) {
  for (int q = pos >> 5; q < DWORDS; ++q) {
    uint32_t word = dword_at(m, q);
    if (q == (pos >> 5)) word &= ~0u << (pos & 31);
    if (word) return q * 32 + __ffs(word) - 1;
  }
  return -1;
}

__device__ __forceinline__ bool has_digi(const uint32_t m[DWORDS], int i) {
  return i < 32 * DWORDS && ((dword_at(m, i >> 5) >> (i & 31)) & 1u);
}

__device__ __forceinline__ void wave_close(const crv_params &p, const Tables &t, Lane &L, const Out &o, Wave &w,
                                           int *flags) {
  const int W = L.W, Sz = p.samples_zs, lmax = W - Sz;
  const bool keep = (float)w.nq * (float)t.vmax >= (float)p.min_voltage;
  w.open = false;
  if (!keep) return;
  const int est = (int)floor((w.t_last - w.t_first + p.single_pe_max_time) / PERIODD);
  int length = est > w.used ? est : w.used;
  if (length > lmax) {
    length = lmax;
    w.overlong = 1;
  }
  while (w.E < length + 2) wave_step(p, L, o, w, true, length);
  if (w.nd == 0 || w.rec < 0) return;  // no digis (the slot stays reserved), or the record pool is full
  const int pedw = adc_f(p, 0.f);  // the rest of the record: pedestal words, whole 8-sample chunks at a time
  int J = w.E;
  for (; J < W && ((W & 7) != 0 || (J & 7) != 0); ++J) put_adc(o, L, w, J, (uint32_t)pedw);
  const uint32_t pp = ((uint32_t)pedw & 0xFFFFu) * 0x10001u;
  for (; J < W; J += 8) *(uint4 *)(o.rec_adc + (int64_t)w.rec * W + J) = make_uint4(pp, pp, pp, pp);
  o.rec_label[w.rec] = L.evkey * CRV_CHAN_MULT + L.ch;
  o.rec_tick[w.rec] = w.kf;
  ++L.used;
  // reco waveforms: runs of continuing digis (start TDC = window tick + first sample), in start order
  const int64_t tdc_base = (int64_t)llrint((w.t_start - L.td) / PERIODD);
  int pos = 0;
  for (;;) {
    const int i0 = next_digi(w.dmask, pos);
    if (i0 < 0) break;
    int last = i0;
    while (has_digi(w.dmask, last + Sz)) last += Sz;
    pos = last + 1;
    const int r = atomicAdd(o.run_count, 1);
    if (r < o.run_cap) {
      o.run_rec[r] = w.rec;
      o.run_i0[r] = i0;
      o.run_n[r] = last + Sz - i0;  // <= W: every digi starts below W - Sz
      o.run_k[r] = L.nrun;
      o.run_c[r] = L.cl;
      o.run_g[r] = L.task;
      o.run_tdc[r] = tdc_base + i0;
    }
    ++L.nrun;
  }
  w.rec = -1;
}

// ---- the waveform stage: one lane per kept charge cluster (kb_wrec) ----

// the end of a pass's descriptors (the reservations of its kb_sipm): drange[1] of the pass, drange[0] of the next
__global__ void kb_dmark(const int32_t *__restrict__ cdesc_count, int32_t *__restrict__ dend) { *dend = *cdesc_count; }

// one lane per kept cluster of the pass, descriptors [drange[0], drange[1]) (grid-stride): the waveform window over
// the cluster's charges -> its ADC record and reco waveforms (run number within the cluster, and its descriptor)
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void kb_wrec(
    crv_params p, Tables t, Derived dv, const uint64_t *__restrict__ task_key, Out o,
    const int32_t *__restrict__ drange, const int *__restrict__ skip, int *__restrict__ flags) {
  __shared__ float tmpl_s[TMPL_LDS];
  if (skip && *skip) return;
  // a full descriptor pool (some SiPMs reserved descriptors they could not write): the call runs again with a
  // larger pool, so none of this pass's descriptors is read
  const int64_t d0 = drange[0], d1 = drange[1];
  if (d1 > o.cdesc_cap) return;
  const int lane = threadIdx.x;
  const DrawR dr{(uint32_t)p.seed, p.evkey_hi};
  for (int i = lane; i < t.ntmpl; i += 64) tmpl_s[i] = t.tmpl_f[i];
  __syncthreads();
  const int64_t lanes = (int64_t)gridDim.x * 64;
  Lane L;
  L.W = p.window_samples;
  L.tmpl = tmpl_s;
  L.used = 0;
  for (int64_t d = d0 + (int64_t)blockIdx.x * 64 + lane; d < d1; d += lanes) {
    const ClDesc cd = o.cdesc[d];
    if (cd.q1 <= cd.q0) {
      o.ncr[d] = 0;
      continue;
    }
    const int64_t g = cd.g;
    const uint64_t key = task_key[g >> 2];
    L.ev = (uint32_t)(key >> 13);
    L.evkey = ((int64_t)p.evkey_hi << 32) | (int64_t)L.ev;
    L.ch = (int)(key & 8191u) * 4 + (int)(g & 3);
    L.task = g;
    L.nrun = 0;
    L.cl = (int32_t)d;
    const uint32_t fk = p.feb_key_roc ? (uint32_t)(t.roc[L.ch] * 1000 + t.feb[L.ch]) : (uint32_t)t.feb[L.ch];
    L.td = (double)u32f(dr(KIND_TDC0, L.ev, fk, 0u, 0u).x) * p.feb_phase_width;
    const double tb0 = o.tbase[g];
    const int64_t lb = cd.lb;
    Wave wv;
    wv.open = false;
    wv.overlong = 0;
    wv.rec = -1;
    int q = cd.q0;
    float4 ck = o.flog[lb + (int64_t)(q >> 1) * 64];
    while (q < cd.q1) {
      const float tqf = (q & 1) ? ck.z : ck.x;
      const double tq = tb0 + (double)tqf;
      float vsum = (q & 1) ? ck.w : ck.y;
      int n = 1;
      for (;;) {
        const int e = q + n;
        if (e >= cd.q1) break;
        if ((e & 1) == 0) ck = o.flog[lb + (int64_t)(e >> 1) * 64];
        if (((e & 1) ? ck.z : ck.x) != tqf) break;
        vsum += (e & 1) ? ck.w : ck.y;
        ++n;
      }
      q += n;
      if (!wv.open) wave_open(L, o, wv, tq);
      wave_add(p, t, L, o, wv, tq, vsum, n, dv.nreach);
    }
    wave_close(p, t, L, o, wv, flags);
    if (wv.rec >= 0) o.rec_label[wv.rec] = -1;  // reserved, never used (the cluster has no digis)
    if (wv.overlong) atomicAdd(&flags[F_OVERLONG], 1);
    o.ncr[d] = L.nrun;
  }
  if (L.used) atomicAdd(o.rec_used, L.used);
}

// per SiPM: the offset of each cluster's reco waveforms within the SiPM (clusters in time order), and the SiPM's total
__global__ void kb_run_order(int64_t L4, const int32_t *__restrict__ Tdev, Out o, const int *__restrict__ skip) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= L4 || (Tdev && g >= 4 * (int64_t)*Tdev) || (skip && *skip)) return;
  const int32_t b0 = o.ncls[g] ? o.cd0[g] : 0, m = o.ncls[g];
  int32_t acc = 0;
  for (int32_t k = 0; k < m; ++k) {
    o.crl[b0 + k] = acc;
    acc += o.ncr[b0 + k];
  }
  o.nrun[g] = acc;
}

// per pool run: its run number within the cluster becomes its run number within the SiPM (the cluster's offset
// added), which places it in the output
__global__ void kb_run_rank(Out o, int64_t nruns, const int32_t *__restrict__ ndev, const int *__restrict__ skip) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((skip && *skip) || r >= (ndev ? min(nruns, (int64_t)*ndev) : nruns)) return;
  o.run_k[r] += o.crl[o.run_c[r]];
}

// kb_fit, one lane per reco waveform of the run pool: pedestal, peaks, Gumbel fits, pulses -> the run's slots; the
// run's place in the output (SiPM's first run + run number) gets its pulse count and pool slot.
// the pulses of pool run r, with the lane's sample scratch y (stride ys)
// Levenberg-Marquardt then Gauss-Newton polish of the Gumbel fit over samples fs..fe of y (stride ys) from the start
// point u (Minuit's sine-mapped parameters; sample q at y[(q - q0) * ys]); u and s end at the best point found
__device__ __forceinline__ void fit_lmgn(const crv_params &p, const float *y, int64_t ys, int q0, int fs, int fe,
                                         const float lo[3], const float hi[3], float u[3], FSum &s) {
  float lam = 1e-3f;
  FSum sn;
  feval(y, ys, q0, fs, fe, lo, hi, u, s);
  for (int it = 0; it < p.lm_iterations; ++it) {
    const double add[3] = {(double)lam * ((double)s.jtj[0] + 1e-12), (double)lam * ((double)s.jtj[3] + 1e-12),
                           (double)lam * ((double)s.jtj[5] + 1e-12)};
    double du[3];
    if (!fsolve(s.jtj, add, s.jtr, du)) du[0] = du[1] = du[2] = 0.0;
    const float un[3] = {u[0] + (float)du[0], u[1] + (float)du[1], u[2] + (float)du[2]};
    // a step that leaves u unchanged (the usual case once the damping is large) evaluates to s itself: not better
    const bool same = un[0] == u[0] && un[1] == u[1] && un[2] == u[2];
    if (!same) feval(y, ys, q0, fs, fe, lo, hi, un, sn);
    const bool better = !same && sn.chi2 < s.chi2;
    if (better) {
      u[0] = un[0];
      u[1] = un[1];
      u[2] = un[2];
      s = sn;
    } else if (lam >= 1e9f) {
      break;  // u, s and lam stay as they are: every remaining iteration would repeat this one
    }
    lam = fminf(fmaxf(better ? lam * 0.3f : lam * 10.f, 1e-9f), 1e9f);
  }
  const double tiny[3] = {1e-12, 1e-12, 1e-12};
  for (int it = 0; it < 8; ++it) {
    double du[3];
    if (!fsolve(s.jtj, tiny, s.jtr, du)) break;
    bool moved = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      du[i] = isfinite(du[i]) ? du[i] : 0.0;
      moved |= du[i] != 0.0;
    }
    const float sc[4] = {1.f, 0.5f, 0.25f, 0.125f};
    bool improved = false;
    for (int k = 0; k < 4 && moved; ++k) {
      const float un[3] = {u[0] + sc[k] * (float)du[0], u[1] + sc[k] * (float)du[1], u[2] + sc[k] * (float)du[2]};
      if (un[0] == u[0] && un[1] == u[1] && un[2] == u[2]) continue;  // evaluates to s itself: not better
      feval(y, ys, q0, fs, fe, lo, hi, un, sn);
      if (sn.chi2 < s.chi2) {
        u[0] = un[0];
        u[1] = un[1];
        u[2] = un[2];
        s = sn;
        moved = false;
        improved = true;
      }
    }
    if (!improved) break;  // u and s unchanged: the remaining iterations would repeat this one
  }
}

__device__ __forceinline__ void fit_run(const crv_params &p, const Tables &t, const uint64_t *__restrict__ task_key,
                                        const int64_t *__restrict__ roff, const Out &o, int64_t r, float *y,
                                        int64_t ys, float *win, int *__restrict__ flags) {
  const int64_t g = o.run_g[r];
  const int ch = (int)(task_key[g >> 2] & 8191u) * 4 + (int)(g & 3);
  const int n = o.run_n[r];
  const int64_t st = o.run_tdc[r];
  const int W = p.window_samples;
  const uint16_t *src = o.rec_adc + (int64_t)o.run_rec[r] * W + o.run_i0[r];
  for (int q = 0; q < n; ++q) y[q * ys] = (float)src[q];
  const bool drop = p.ignore_channels && (t.status[ch] & 0b010) != 0;
  const float ped_db = (float)t.pedestal[ch];
  const bool und = ped_db - y[0] > (float)p.pedestal_undershoot_threshold;
  const float ped = und ? y[0] : ped_db;
  for (int i = 0; i < n; ++i) y[i * ys] -= ped;
  int np = 0;
  int startb = 0;
  bool failed_before = false;
  for (int seq = 0;; ++seq) {
    bool found = false, rising = false;
    int ps = 0, pe = 0;
    for (int b = startb + 1; b < n; ++b) {
      const float prev = y[(b - 1) * ys], cur = y[b * ys];
      if (prev < cur) {
        rising = true;
        ps = pe = b;
      } else if (prev == cur && rising) {
        pe = b;
      } else if (prev > cur && rising) {
        if (y[pe * ys] > (float)p.min_adc_difference) {
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
      const int i = ps - k;
      const bool cond = i >= 0 && y[i * ys] <= y[(i + 1) * ys];
      if (cond) fs = i;
      go = cond && i != 0;
    }
    go = true;
    for (int k = 1; k <= p.samples_after && go; ++k) {
      const int i = pe + k;
      const bool cond = i < n && y[(i - 1) * ys] >= y[i * ys];
      if (cond) fe = i;
      go = cond;
    }
    if (ps - fs > 1) fs += 1;
    if (fe - pe > 1) fe -= 1;
    const float yps = y[ps * ys];
    const float t_ps = (float)ps * PERIODF, t_pe = (float)pe * PERIODF;  // relative to the waveform start
    const float t_peak = 0.5f * (t_ps + t_pe);
    const float ye = yps * 2.71828183f;
    const float lo[3] = {ye * (float)p.min_pulse_height_ratio, t_ps - (float)p.max_time_difference, (float)p.min_beta};
    const float hi[3] = {ye * (float)p.max_pulse_height_ratio, t_pe + (float)p.max_time_difference, (float)p.max_beta};
    const float p0[3] = {ye, t_peak, (float)p.default_beta};
    float u[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) u[i] = asinf(fminf(fmaxf(2.f * (p0[i] - lo[i]) / (hi[i] - lo[i]) - 1.f, -1.f), 1.f));
    FSum s;
    const bool inw = fe - fs + 1 <= FWIN;  // the fit window in LDS ([sample][lane]): every evaluation reads it
    if (inw)
      for (int q = fs; q <= fe; ++q) win[(q - fs) * 64] = y[q * ys];
    fit_lmgn(p, inw ? win : y, inw ? 64 : ys, inw ? fs : 0, fs, fe, lo, hi, u, s);
    const double tiny[3] = {1e-12, 1e-12, 1e-12};
    double step[3];
    const bool ok = fsolve(s.jtj, tiny, s.jtr, step);
    const double edm = 0.5 * ((double)s.jtr[0] * step[0] + (double)s.jtr[1] * step[1] + (double)s.jtr[2] * step[2]);
    float emin, emax;
    feig(s.jtj, emin, emax);
    const bool hess_ok = ok && emin > 1e-12f * fmaxf(emax, 1e-30f);
    const bool conv = ok && isfinite(edm) && edm <= 1e-3 * fmax((double)s.chi2, 1.0);
    float par[3], dp[3];
    fparams(lo, hi, u, par, dp);
    bool near = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float span = hi[i] - lo[i];
      near |= (par[i] - lo[i]) / span < (float)p.limit_tolerance || (hi[i] - par[i]) / span < (float)p.limit_tolerance;
    }
    const bool failed = !conv || !hess_ok || near;
    const float A = par[0], mu = par[1], bb = par[2];
    const int ndf = (fe - fs + 1) - 3;
    const bool fseq = failed || failed_before;
    if (!drop) {
      if (np < o.pcap) {
        const int64_t q = r * o.pcap + np;
        const double t0 = (double)st * PERIODD;
        Slot sv;
        sv.tp = t0 + (double)(failed ? t_peak : mu);
        sv.mu = t0 + (double)mu;
        sv.pe = (failed ? ye * (float)p.default_beta : A * bb) / (float)t.area[ch];
        sv.height = failed ? yps : A / 2.71828183f;
        sv.beta = failed ? (float)p.default_beta : bb;
        sv.chi2 = (failed || ndf <= 0) ? -1.f : s.chi2 / (float)max(ndf, 1);
        sv.ped = ped;
        sv.fa = A;
        sv.fb = bb;
        sv.sf = (uint32_t)seq | (fseq ? 1u << 16 : 0u) | (ndf <= 0 ? 1u << 17 : 0u) | (!und ? 1u << 18 : 0u);
        o.slot[q] = sv;
      }
      ++np;
    }
    failed_before = fseq;
    const float ib = 1.f / bb;
    for (int i = 0; i < n; ++i) {
      const float z = ((float)i * PERIODF - mu) * ib;
      const float sub = A * __expf(-z - __expf(-z));
      y[i * ys] -= sub >= 0.f ? floorf(sub + 0.5f) : -floorf(-sub + 0.5f);
    }
    startb = fe;
  }
  if (np > o.pcap) atomicAdd(&flags[F_PSLOT], 1);
  const int64_t fr = roff[g] + o.run_k[r];
  o.npr[fr] = min(np, o.pcap);
  o.rmap[fr] = (int32_t)r;
}

__global__ __launch_bounds__(64) void kb_fit(crv_params p, Tables t, const uint64_t *__restrict__ task_key,
                                             const int64_t *__restrict__ roff, Out o, int64_t nruns,
                                             int *__restrict__ flags, const int32_t *__restrict__ ndev = nullptr,
                                             const int *__restrict__ skip = nullptr) {
  __shared__ float win[FWIN * 64];
  if (skip && *skip) return;
  if (ndev) nruns = min(nruns, (int64_t)*ndev);  // fixed-capacity calls: the runs of the call
  const int64_t lanes = (int64_t)gridDim.x * blockDim.x, gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = gid; r < nruns; r += lanes)
    fit_run(p, t, task_key, roff, o, r, o.yscr + gid, lanes, win + threadIdx.x, flags);
}

// fired-pixel log capacity per lane of each group of 64 SiPMs in queue order, from the largest photon count of the
// group (3 factor / 4 of it + 16; factor doubles on every re-run); the group takes 64 times that
// log capacity per group of 64 queue positions: one wavefront per group, each lane one SiPM (coalesced queue reads),
// the group's largest photon count by a wavefront max
__global__ void kb_gcaps(const int64_t *__restrict__ nph, const int32_t *__restrict__ order, int64_t L4,
                         int64_t *__restrict__ gcap, int64_t *__restrict__ gsize, int factor,
                         const int32_t *__restrict__ Tdev = nullptr) {
  const int lane = threadIdx.x & 63;
  const int64_t grp = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (grp >= (L4 + 63) / 64) return;  // (whole wavefronts: every lane of a group returns together)
  if (Tdev && grp * 64 >= 4 * (int64_t)*Tdev) {  // fixed-capacity calls: groups past the call hold nothing
    if (lane == 0) {
      gcap[grp] = 0;
      gsize[grp] = 0;
    }
    return;
  }
  if (Tdev) L4 = min(L4, 4 * (int64_t)*Tdev);
  const int64_t gq = grp * 64 + lane;
  int64_t m = 0;
  if (gq < L4) m = nph[order ? order[gq] : gq] * 3 * factor / 4 + 16;  // fired pixels: about 42 % of the photons
  for (int o = 32; o > 0; o >>= 1) m = max(m, (int64_t)__shfl_xor(m, o));
  if (lane == 0) {
    m = (m + 1) & ~(int64_t)1;
    gcap[grp] = m;
    gsize[grp] = m / 2 * 64;  // chunks of two entries
  }
}

__global__ void kb_zero(int *__restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void kb_widen(const int32_t *__restrict__ in, int64_t *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// the pulse table in run order (= (event, channel, time) order), one lane per run
__global__ void kb_pulses_out(crv_params p, Tables t, int64_t nruns, const uint64_t *__restrict__ task_key,
                              const int64_t *__restrict__ poff, Out o, crv_pulses P, int32_t *__restrict__ pbar,
                              int32_t *__restrict__ psipm, const int32_t *__restrict__ ndev = nullptr,
                              const int *__restrict__ skip = nullptr) {
  const int64_t fr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (skip && *skip) return;
  if (ndev) nruns = min(nruns, (int64_t)*ndev);
  if (fr >= nruns) return;
  const int64_t dst = poff[fr], n = poff[fr + 1] - dst;
  if (n == 0) return;
  const int64_t r = o.rmap[fr];
  const int64_t g = o.run_g[r];
  const uint64_t key = task_key[g >> 2];
  const int ch = (int)(key & 8191u) * 4 + (int)(g & 3);
  const int64_t evkey = ((int64_t)p.evkey_hi << 32) | (int64_t)(key >> 13);
  const double toff = p.use_time_offset_db ? t.toff[ch] * p.time_offset_scale : 0.0;
  const double hdb = t.height[ch];
  for (int64_t k = 0; k < n; ++k) {
    const int64_t b = dst + k;
    const Slot sv = o.slot[r * o.pcap + k];
    const uint32_t sf = sv.sf;
    const double tp = sv.tp, beta = sv.beta, height = sv.height;
    P.wf[b] = fr;
    P.seq[b] = sf & 0xFFFFu;
    P.ev[b] = evkey;
    P.chan[b] = ch;
    P.pe[b] = sv.pe;
    P.peph[b] = height / hdb;
    P.time[b] = tp + toff;
    P.height[b] = height;
    P.beta[b] = beta;
    P.chi2[b] = sv.chi2;
    P.letime[b] = tp - p.le_time_factor * beta;
    P.ped[b] = sv.ped;
    P.fit_a[b] = sv.fa;
    P.fit_mu[b] = sv.mu;
    P.fit_b[b] = sv.fb;
    P.failed[b] = (sf >> 16) & 1u;
    P.zerondf[b] = (sf >> 17) & 1u;
    P.pedfromdb[b] = (sf >> 18) & 1u;
    pbar[b] = ch / 4;
    psipm[b] = ch % 4;
  }
}

int resident_blocks(Ctx *c, const void *kernel) {
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
  return per_cu * cus;
}

}  // namespace

// ---- fixed-capacity calls (captured into a graph): padding and capacity checks ----

// SiPM positions past the call's 4 T: a sort key after every real one (their stable order keeps them last), no
// photons, no fired pixels, no runs
__global__ void kb_pad_sipms(int64_t L4cap, const int32_t *__restrict__ Tdev, uint32_t pad_key,
                             uint32_t *__restrict__ skey, int32_t *__restrict__ sidx, int64_t *__restrict__ sipm_nph,
                             int32_t *__restrict__ nfired, int32_t *__restrict__ nrun) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= L4cap || g < 4 * (int64_t)*Tdev) return;
  skey[g] = pad_key;
  sidx[g] = (int32_t)g;
  sipm_nph[g] = 0;
  nfired[g] = 0;
  nrun[g] = 0;
}

// capacity checks: `skip` set (with the first reason) when the call's log, pools, pulse slots or pulses exceed them
__global__ void kb_check_log(const int64_t *__restrict__ total, int64_t cap, int *__restrict__ skip) {
  if (*total > cap && atomicCAS(skip, 0, 1) == 0) skip[1] = 1;
}
__global__ void kb_check_wave(const int *__restrict__ flags, const int32_t *__restrict__ counters, int64_t rec_cap,
                              int64_t run_cap, int64_t desc_cap, int *__restrict__ skip) {
  // a log overflow of kb_sipm shows in F_OVERFLOW
  if (flags[F_OVERFLOW] && atomicCAS(skip, 0, 1) == 0) skip[1] = 1;
  if ((counters[C_REC] > rec_cap || counters[C_RUN] > run_cap || counters[C_DESC] > desc_cap) &&
      atomicCAS(skip, 0, 1) == 0)
    skip[1] = 2;
}
__global__ void kb_check_fit(const int *__restrict__ flags, const int64_t *__restrict__ np, int64_t np_cap,
                             int *__restrict__ skip) {
  if (flags[F_PSLOT] && atomicCAS(skip, 0, 1) == 0) skip[1] = 3;
  if (*np > np_cap && atomicCAS(skip, 0, 1) == 0) skip[1] = 4;
}

__global__ void kb_zero32(int32_t *__restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// pulse rows past the call's pulses (all of them when the call was skipped): an event key after every real one, so
// the finder's grouping sees them as one trailing event of no pulses
__global__ void kb_pad_pulses(int64_t np_cap, const int64_t *__restrict__ np, const int *__restrict__ skip,
                              int64_t *__restrict__ ev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np_cap) return;
  if (*skip || i >= *np) ev[i] = INT64_MAX;
}

// the call's sizes for the host: pulses, photons, fired pixels, records, runs (clusters and members are written by
// the finder's part)
__global__ void kb_sizes(const int64_t *__restrict__ np, const int64_t *__restrict__ sums,
                         const int32_t *__restrict__ counters, const int *__restrict__ skip, int64_t *__restrict__ sz) {
  sz[0] = *skip ? 0 : *np;
  sz[3] = sums[1];
  sz[4] = sums[0];
  sz[5] = counters[C_USED];  // records with digis
  sz[6] = counters[C_RUN];
}

// ---- host side ----
int batched_stage(Ctx *c, const crv_steps *steps, crv_batch_opts *opt, crv_timing *timing, hipStream_t s) {
  c->ws_spare.clear();
  const crv_params &p = c->p;
  const int W = p.window_samples, Sz = p.samples_zs;
  if (W > 32 * DWORDS || Sz < 1 || W <= Sz)
    return set_error(c, CRV_E_ARG, "batched chain: window_samples must be in (samples_zs, 192]");
  if (!c->t.tmpl_f) return set_error(c, CRV_E_STATE, "template not uploaded");
  if (c->t.ntmpl > TMPL_LDS)
    return set_error(c, CRV_E_ARG, "batched chain: single-PE template longer than " + std::to_string(TMPL_LDS) +
                                       " samples");
  if (c->d.nreach > RINGB) return set_error(c, CRV_E_ARG, "batched chain: single-PE template reaches over 16 samples");
  const int tile = opt->tile > 0 ? opt->tile : 64;
  if (tile % 64) return set_error(c, CRV_E_ARG, "batched chain: tile must be a multiple of 64");
  if (opt->sort_tasks < 0 || opt->sort_tasks > 3)
    return set_error(c, CRV_E_ARG, "batched chain: sort_tasks must be 0, 1, 2 or 3");
  const int grid_s = opt->grid > 0 ? opt->grid : resident_blocks(c, (const void *)kb_sipm);
  const int grid_r = opt->grid > 0 ? opt->grid : resident_blocks(c, (const void *)kb_wrec);
  hipEvent_t *ev = c->ev, *kev = c->kev;
  CRV_HIP(c, hipEventRecord(ev[0], s));
  StepGroups g;
  int rc = group_steps(c, steps, g, s);
  if (rc) return rc;
  const int64_t T = g.T, L4 = T * 4, n = g.n_steps;
  if (L4 > INT32_MAX / 2) return set_error(c, CRV_E_ARG, "batched chain: too many hit counters in one call");
  int32_t *Tdev = (int32_t *)ws(c, 500, 8);
  int32_t *nph = (int32_t *)ws(c, 501, n * 16 + 16);
  int64_t *sipm_nph = (int64_t *)ws(c, 502, (L4 + 1) * 8);
  const int64_t ngrp = (L4 + 63) / 64;
  int64_t *gcap = (int64_t *)ws(c, 503, (ngrp + 1) * 8), *gsize = (int64_t *)ws(c, 518, (ngrp + 1) * 8);
  int64_t *goff = (int64_t *)ws(c, 504, (ngrp + 2) * 8);
  uint32_t *skey = (uint32_t *)ws(c, 505, L4 * 4 + 8), *skey2 = (uint32_t *)ws(c, 506, L4 * 4 + 8);
  int32_t *sidx = (int32_t *)ws(c, 507, L4 * 4 + 8), *order = (int32_t *)ws(c, 508, L4 * 4 + 8);
  int32_t *nfired = (int32_t *)ws(c, 512, L4 * 4 + 8);
  int *counters = (int *)ws(c, 540, C_INTS * 4);
  Out o{};
  o.xstride = (int64_t)grid_s * 64;
  o.xcur = (uint16_t *)ws(c, 543, (size_t)o.xstride * CMAXB * 2);
  o.xcand = (uint16_t *)ws(c, 544, (size_t)o.xstride * CMAXB * 2);
  o.xmult = (uint8_t *)ws(c, 545, (size_t)o.xstride * CMAXB);
  o.lastt = (float *)ws(c, 539, (size_t)o.xstride * CRV_NPIX * CRV_NPIX * 4);
  o.gS = (float *)ws(c, 534, n * 16 + 16);
  o.gHead = (double *)ws(c, 535, n * 32 + 16);
  o.gTb = (double *)ws(c, 536, n * 32 + 16);
  o.gRJ = (uint64_t *)ws(c, 537, n * 32 + 16);
  o.nrun = (int32_t *)ws(c, 513, L4 * 4 + 8);
  o.cd0 = (int32_t *)ws(c, 521, L4 * 4 + 8);
  o.ncls = (int32_t *)ws(c, 522, L4 * 4 + 8);
  if (!o.cd0 || !o.ncls) return ws_fail(c);
  if (!Tdev || !nph || !sipm_nph || !gcap || !gsize || !goff || !skey || !skey2 || !sidx || !order || !nfired || !counters ||
      !o.xcur || !o.xcand || !o.xmult || !o.lastt || !o.gS || !o.gHead || !o.gTb || !o.gRJ || !o.nrun)
    return ws_fail(c);
  o.rec_count = counters + C_REC;
  o.run_count = counters + C_RUN;
  o.cdesc_count = counters + C_DESC;
  o.rec_used = counters + C_USED;
  const int32_t T32 = (int32_t)T;
  CRV_HIP(c, hipMemcpyAsync(Tdev, &T32, 4, hipMemcpyHostToDevice, s));
  kb_zero<<<1, 64, 0, s>>>(c->flags, F_NFLAGS);
  const BStep st{g.edep, g.z, g.t};
  if (L4)
    kb_count<<<nblocks(L4, 256), 256, 0, s>>>(p, c->t, c->d, st, Tdev, g.task_key, g.task_off, g.perm, g.evfirst,
                                              nph, sipm_nph, skey, sidx, opt->sort_tasks, c->flags);
  CRV_HIP(c, hipGetLastError());
  const bool sorted = opt->sort_tasks != 0 && L4 > 0;
  const int sbits = sort_key_bits(opt->sort_tasks);
  if (sorted) {
    size_t tb = 0;
    CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, skey, skey2, sidx, order, (int)L4, 0, sbits, s));
    void *tmp = ws(c, 509, tb);
    if (!tmp) return ws_fail(c);
    CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, skey, skey2, sidx, order, (int)L4, 0, sbits, s));
  }
  CRV_HIP(c, hipEventRecord(ev[1], s));
  int hf[F_NFLAGS];
  int32_t cnt[3] = {0, 0, 0};  // records, runs, cluster descriptors
  auto read_counts = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
      e = hipMemcpyAsync(cnt + k, counters + C_REC + k * CL, 4, hipMemcpyDeviceToHost, s);
    return e;
  };
  float ms_sipm = 0.f, ms_wave = 0.f;  // summed over the passes of the last attempt
  const int32_t *qorder = sorted ? order : nullptr;
  auto kernel_ms = [&](int k) {
    float e = 0.f;
    (void)hipEventElapsedTime(&e, kev[2 * k], kev[2 * k + 1]);
    return e;
  };
  // kb_sipm over queue positions [t0, t1): 0, 1 when a fired-pixel log overflowed, or an error code
  // (and the end of its descriptors in dpass[pi + 1])
  int32_t *dpass = nullptr;  // per pass its first descriptor, then the end of the last pass's
  auto run_sipm = [&](int64_t t0, int64_t t1, int pi) -> int {
    kb_zero<<<1, 64, 0, s>>>(counters, 1);
    kb_zero<<<1, 64, 0, s>>>(c->flags + F_OVERFLOW, 1);
    CRV_HIP(c, hipEventRecord(kev[0], s));
    kb_sipm<<<grid_s, 64, 0, s>>>(p, c->t, c->d, st, g.task_key, g.task_off, g.perm, g.evfirst, nph, o,
                                  Queue{counters, tile, t1, qorder, t0}, nfired, c->flags);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipEventRecord(kev[1], s));
    kb_dmark<<<1, 1, 0, s>>>(o.cdesc_count, dpass + pi + 1);
    CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    CRV_HIP(c, hipStreamSynchronize(s));
    ms_sipm += kernel_ms(0);
    if (hf[F_CASCADE])
      return set_error(c, CRV_E_OVERFLOW,
                       "batched chain: cross-talk cascade wider than 16 pixels or deeper than 31 generations");
    if (hf[F_COUNTER]) return set_error(c, CRV_E_OVERFLOW, "batched chain: more than 65535 photoelectrons from one step");
    return hf[F_OVERFLOW] ? 1 : 0;
  };
  auto run_wave = [&](int pi) -> int {
    CRV_HIP(c, hipEventRecord(kev[2], s));
    kb_wrec<<<grid_r, 64, 0, s>>>(p, c->t, c->d, g.task_key, o, dpass + pi, nullptr, c->flags);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipEventRecord(kev[3], s));
    CRV_HIP(c, hipStreamSynchronize(s));
    ms_wave += kernel_ms(1);
    return 0;
  };
  // record and run pools at their current capacities
  auto pools = [&]() -> int {
    o.rec_cap = std::max<int64_t>(c->b_rec_cap, L4 + 1024);
    o.run_cap = std::max<int64_t>(c->b_run_cap, L4 + 1024);
    o.rec_adc = (uint16_t *)ws(c, 570, o.rec_cap * W * 2);
    o.rec_label = (int64_t *)ws(c, 571, o.rec_cap * 8);
    o.rec_tick = (int32_t *)ws(c, 572, o.rec_cap * 4);
    o.run_rec = (int32_t *)ws(c, 573, o.run_cap * 4);
    o.run_i0 = (int32_t *)ws(c, 574, o.run_cap * 4);
    o.run_n = (int32_t *)ws(c, 575, o.run_cap * 4);
    o.run_k = (int32_t *)ws(c, 576, o.run_cap * 4);
    o.run_g = (int64_t *)ws(c, 577, o.run_cap * 8);
    o.run_tdc = (int64_t *)ws(c, 578, o.run_cap * 8);
    o.run_c = (int32_t *)ws(c, 523, o.run_cap * 4);
    o.cdesc_cap = std::max<int64_t>(c->b_desc_cap, L4 + 1024);
    o.cdesc = (ClDesc *)ws(c, 524, o.cdesc_cap * sizeof(ClDesc));
    o.ncr = (int32_t *)ws(c, 525, o.cdesc_cap * 4);
    o.crl = (int32_t *)ws(c, 526, o.cdesc_cap * 4);
    if (!o.rec_adc || !o.rec_label || !o.rec_tick || !o.run_rec || !o.run_i0 || !o.run_n || !o.run_k || !o.run_g ||
        !o.run_tdc || !o.run_c || !o.cdesc || !o.ncr || !o.crl)
      return ws_fail(c);
    return 0;
  };
  // Fired-pixel log budget (bytes): the caller's, else half of the device memory free now plus what
  // the log already holds. When the whole call's log exceeds it, the queue is cut into passes of whole groups whose
  // logs fit, and each pass runs kb_sipm then kb_wrec on its SiPMs (records and runs accumulate).
  size_t budget = opt->log_budget > 0 ? (size_t)opt->log_budget : 0;
  if (!budget) {
    size_t fr = 0, total = 0;
    CRV_HIP(c, hipMemGetInfo(&fr, &total));
    const size_t held = ((int)c->ws_size.size() > 538 ? c->ws_size[538] : 0);
    budget = (fr + held) / 2;
  }
  std::vector<int64_t> hoff, cuts;  // group offsets on the host; first group of each pass, then ngrp
  int64_t nlog = 0;
  int npass = 1;
  // phases 1 and 2: photons -> fired-pixel logs -> records and reco waveforms. A log overflow grows the logs and runs
  // everything again; a pool overflow grows the pools and runs the waveform kernel again (and kb_sipm, when passes
  // reuse the log or the descriptor pool was full).
  for (int attempt = 0, factor = c->b_log_factor;; ++attempt, factor *= 2) {
    if (attempt > 8) return set_error(c, CRV_E_OVERFLOW, "batched chain: fired-pixel log still too short");
    if (L4)
      kb_gcaps<<<nblocks(ngrp * 64, 256), 256, 0, s>>>(sipm_nph, qorder, L4, gcap, gsize, factor);
    if ((rc = scan_i64(c, gsize, goff, ngrp, 510, s))) return rc;
    if (attempt == 0) CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    if ((rc = read_i64(c, goff + ngrp, &nlog, s))) return rc;  // (synchronises the stream)
    if (attempt == 0 && hf[F_BADCHAN])
      return set_error(c, CRV_E_ARG, "batched chain: channel beyond the uploaded conditions");
    int64_t log_chunks = nlog;  // goff and nlog count 16-byte chunks
    cuts.assign({0, ngrp});
    bool pipe = opt->pipeline >= 2 && ngrp >= 2;  // two log buffers, each pass within half the budget
    if (pipe || (size_t)nlog * 16 > budget) {
      hoff.resize(ngrp + 1);
      CRV_HIP(c, hipMemcpyAsync(hoff.data(), goff, (ngrp + 1) * 8, hipMemcpyDeviceToHost, s));
      CRV_HIP(c, hipStreamSynchronize(s));
      int64_t lim = (int64_t)(budget / 16), gmax = 0;
      for (int64_t gi = 0; gi < ngrp; ++gi) gmax = std::max<int64_t>(gmax, hoff[gi + 1] - hoff[gi]);
      if (pipe) {
        // a pass holds at least the largest group; when two buffers of that size exceed the budget, the passes run
        // one after the other in one buffer (results are the same either way)
        if (gmax > lim / 2) pipe = false;
        else lim = std::min<int64_t>(lim / 2, std::max<int64_t>(gmax, (nlog + opt->pipeline - 1) / opt->pipeline));
      }
      cuts.assign(1, 0);
      log_chunks = 0;
      for (int64_t g0 = 0; g0 < ngrp;) {
        const int64_t g1 = std::upper_bound(hoff.begin() + g0 + 1, hoff.end(), hoff[g0] + lim) - hoff.begin() - 1;
        if (g1 <= g0)
          return set_error(c, CRV_E_ARG, "batched chain: the fired-pixel log of one group of 64 SiPMs (" +
                                             std::to_string((hoff[g0 + 1] - hoff[g0]) * 16 >> 20) +
                                             " MiB) exceeds the log budget");
        log_chunks = std::max(log_chunks, hoff[g1] - hoff[g0]);
        cuts.push_back(g1);
        g0 = g1;
      }
    }
    npass = (int)cuts.size() - 1;
    c->b_last_log = nlog;
    o.goff = goff;
    o.gcap = gcap;
    o.flog = (float4 *)ws(c, 538, log_chunks * 16 + 16);
    o.tbase = (double *)ws(c, 519, L4 * 8 + 8);
    dpass = (int32_t *)ws(c, 527, (size_t)(npass + 2) * 4);
    if (!o.flog || !o.tbase || !dpass) return ws_fail(c);
    float4 *flog2 = nullptr;
    if (pipe && npass > 1) {
      flog2 = (float4 *)ws(c, 547, log_chunks * 16 + 16);
      if (!flog2) return ws_fail(c);
    }
    bool grow_log = false;
    if (flog2) {  // pipelined passes: kb_sipm of pass i + 1 on `s` beside the waveform kernels of pass i (side stream)
      const hipStream_t s2 = c->aux[0];
      while ((int)c->pev.size() < 6 * npass + 2) {
        hipEvent_t e;
        CRV_HIP(c, hipEventCreate(&e));
        c->pev.push_back(e);
      }
      hipEvent_t *S = c->pev.data(), *Wv = S + npass, *t0s = S + 2 * npass, *t1s = S + 3 * npass,
                 *t0w = S + 4 * npass, *t1w = S + 5 * npass, fk = S[6 * npass], jn = S[6 * npass + 1];
      for (int pa = 0;; ++pa) {
        if ((rc = pools())) return rc;
        kb_zero<<<1, 64, 0, s>>>(counters + C_REC, C_USED + 1 - C_REC);
        kb_zero<<<1, 64, 0, s>>>(dpass, 1);
        kb_zero<<<1, 64, 0, s>>>(c->flags + F_OVERFLOW, 1);
        CRV_HIP(c, hipEventRecord(fk, s));
        CRV_HIP(c, hipStreamWaitEvent(s2, fk, 0));
        for (int pi = 0; pi < npass; ++pi) {
          const int64_t q0 = cuts[pi] * 64, q1 = std::min<int64_t>(cuts[pi + 1] * 64, L4);
          Out op = o;
          op.flog = (pi & 1) ? flog2 : o.flog;
          op.log0 = hoff[cuts[pi]];
          if (pi >= 2) CRV_HIP(c, hipStreamWaitEvent(s, Wv[pi - 2], 0));  // its log buffer has been read
          kb_zero<<<1, 64, 0, s>>>(counters, 1);
          CRV_HIP(c, hipEventRecord(t0s[pi], s));
          kb_sipm<<<grid_s, 64, 0, s>>>(p, c->t, c->d, st, g.task_key, g.task_off, g.perm, g.evfirst, nph, op,
                                        Queue{counters, tile, q1, qorder, q0}, nfired, c->flags);
          CRV_HIP(c, hipGetLastError());
          CRV_HIP(c, hipEventRecord(t1s[pi], s));
          kb_dmark<<<1, 1, 0, s>>>(o.cdesc_count, dpass + pi + 1);
          CRV_HIP(c, hipEventRecord(S[pi], s));
          CRV_HIP(c, hipStreamWaitEvent(s2, S[pi], 0));
          CRV_HIP(c, hipEventRecord(t0w[pi], s2));
          kb_wrec<<<grid_r, 64, 0, s2>>>(p, c->t, c->d, g.task_key, op, dpass + pi, nullptr, c->flags);
          CRV_HIP(c, hipGetLastError());
          CRV_HIP(c, hipEventRecord(t1w[pi], s2));
          CRV_HIP(c, hipEventRecord(Wv[pi], s2));
        }
        CRV_HIP(c, hipEventRecord(jn, s2));
        CRV_HIP(c, hipStreamWaitEvent(s, jn, 0));
        CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
        CRV_HIP(c, read_counts());
        CRV_HIP(c, hipStreamSynchronize(s));
        ms_sipm = ms_wave = 0.f;
        for (int pi = 0; pi < npass; ++pi) {
          float e1 = 0.f, e2 = 0.f;
          (void)hipEventElapsedTime(&e1, t0s[pi], t1s[pi]);
          (void)hipEventElapsedTime(&e2, t0w[pi], t1w[pi]);
          ms_sipm += e1;
          ms_wave += e2;
        }
        if (hf[F_CASCADE])
          return set_error(c, CRV_E_OVERFLOW,
                           "batched chain: cross-talk cascade wider than 16 pixels or deeper than 31 generations");
        if (hf[F_COUNTER])
          return set_error(c, CRV_E_OVERFLOW, "batched chain: more than 65535 photoelectrons from one step");
        if (hf[F_OVERFLOW]) {
          grow_log = true;
          break;
        }
        if (cnt[0] <= o.rec_cap && cnt[1] <= o.run_cap && cnt[2] <= o.cdesc_cap) break;
        if (pa >= 3) return set_error(c, CRV_E_OVERFLOW, "batched chain: record or run pool still too small");
        c->b_rec_cap = std::max<int64_t>(o.rec_cap, (int64_t)cnt[0] + cnt[0] / 4 + 1024);
        c->b_run_cap = std::max<int64_t>(o.run_cap, (int64_t)cnt[1] + cnt[1] / 4 + 1024);
        c->b_desc_cap = std::max<int64_t>(o.cdesc_cap, (int64_t)cnt[2] + cnt[2] / 4 + 1024);
      }
      if (!grow_log) {
        c->b_log_factor = factor;
        break;
      }
      continue;
    }
    for (int pa = 0, desc_full = 0;; ++pa) {
      if ((rc = pools())) return rc;
      // kb_sipm runs again when the passes reuse the log or its descriptors did not fit; otherwise only the
      // waveform kernel runs again, over the descriptors of the last run
      const bool sipm = pa == 0 || npass > 1 || desc_full;
      kb_zero<<<1, 64, 0, s>>>(counters + C_REC, 1);
      kb_zero<<<1, 64, 0, s>>>(counters + C_RUN, 1);
      kb_zero<<<1, 64, 0, s>>>(counters + C_USED, 1);
      if (sipm) {
        kb_zero<<<1, 64, 0, s>>>(counters + C_DESC, 1);
        kb_zero<<<1, 64, 0, s>>>(dpass, 1);
        ms_sipm = 0.f;
      }
      ms_wave = 0.f;
      for (int pi = 0; pi < npass && L4; ++pi) {
        const int64_t t0 = cuts[pi] * 64, t1 = std::min<int64_t>(cuts[pi + 1] * 64, L4);
        o.log0 = npass > 1 ? hoff[cuts[pi]] : 0;
        if (sipm) {
          if ((rc = run_sipm(t0, t1, pi)) < 0) return rc;
          if (rc) {
            grow_log = true;
            break;
          }
        }
        if ((rc = run_wave(pi))) return rc;
      }
      if (grow_log) break;
      CRV_HIP(c, read_counts());
      CRV_HIP(c, hipStreamSynchronize(s));
      if (cnt[0] <= o.rec_cap && cnt[1] <= o.run_cap && cnt[2] <= o.cdesc_cap) break;
      if (pa >= 3) return set_error(c, CRV_E_OVERFLOW, "batched chain: record or run pool still too small");
      desc_full = cnt[2] > o.cdesc_cap;
      c->b_rec_cap = std::max<int64_t>(o.rec_cap, (int64_t)cnt[0] + cnt[0] / 4 + 1024);
      c->b_run_cap = std::max<int64_t>(o.run_cap, (int64_t)cnt[1] + cnt[1] / 4 + 1024);
      c->b_desc_cap = std::max<int64_t>(o.cdesc_cap, (int64_t)cnt[2] + cnt[2] / 4 + 1024);
    }
    if (!grow_log) {
      c->b_log_factor = factor;
      break;
    }
  }
  // the photon and waveform kernels are done: their fired-pixel logs, last-fire times, photon streams and cascade
  // lists may be freed if the fit or the finder needs the room (they are allocated again by the next call)
  c->ws_spare = {534, 535, 536, 537, 538, 539, 543, 544, 545, 547};
  // photons drawn and pixels fired (diagnostics)
  int64_t *sums = (int64_t *)ws(c, 515, 16);
  int64_t *fired64 = (int64_t *)ws(c, 516, L4 * 8 + 8);
  if (!sums || !fired64) return ws_fail(c);
  int64_t hsum[2] = {0, 0};
  if (L4) {
    kb_widen<<<nblocks(L4, 256), 256, 0, s>>>(nfired, fired64, L4);
    size_t tb = 0;
    CRV_HIP(c, hipcub::DeviceReduce::Sum(nullptr, tb, fired64, sums, (int)L4, s));
    void *tmp = ws(c, 517, tb);
    if (!tmp) return ws_fail(c);
    CRV_HIP(c, hipcub::DeviceReduce::Sum(tmp, tb, fired64, sums, (int)L4, s));
    CRV_HIP(c, hipcub::DeviceReduce::Sum(tmp, tb, sipm_nph, sums + 1, (int)L4, s));
    CRV_HIP(c, hipMemcpyAsync(hsum, sums, 16, hipMemcpyDeviceToHost, s));
  }
  const int64_t nrec = cnt[0], nruns = cnt[1];
  c->b_last_desc = cnt[2];
  int32_t nused = 0;
  CRV_HIP(c, hipMemcpyAsync(&nused, o.rec_used, 4, hipMemcpyDeviceToHost, s));
  // the runs' places in the output: each SiPM's runs in run-number order
  int64_t *nrun64 = (int64_t *)ws(c, 579, (L4 + 1) * 8), *roff = (int64_t *)ws(c, 580, (L4 + 2) * 8);
  if (!nrun64 || !roff) return ws_fail(c);
  if (L4) kb_run_order<<<nblocks(L4, 256), 256, 0, s>>>(L4, nullptr, o, nullptr);
  if (nruns) kb_run_rank<<<nblocks(nruns, 256), 256, 0, s>>>(o, nruns, nullptr, nullptr);
  if (L4) kb_widen<<<nblocks(L4, 256), 256, 0, s>>>(o.nrun, nrun64, L4);
  if ((rc = scan_i64(c, nrun64, roff, L4, 514, s))) return rc;
  // phase 3: pulse fits (slots grown and the fits run again when a waveform has more pulses)
  const int fit_blocks = (int)std::min<int64_t>(nblocks(nruns, 64), 8192);  // grid-stride; scratch per lane
  o.yscr = (float *)ws(c, 581, (size_t)fit_blocks * 64 * W * 4 + 8);
  o.npr = (int32_t *)ws(c, 582, nruns * 4 + 8);
  o.rmap = (int32_t *)ws(c, 583, nruns * 4 + 8);
  if (!o.yscr || !o.npr || !o.rmap) return ws_fail(c);
  // the caller's slot count, else the count the context's earlier calls grew to (kept like the log factor, so a
  // workload whose waveforms hold more pulses fits its slots from the second call on)
  o.pcap = opt->pulse_slots > 0 ? opt->pulse_slots : std::max(PSLOTS, c->b_pcap);
  for (int attempt = 0;; ++attempt) {
    const int64_t ns = nruns * o.pcap + 8;
    o.slot = (Slot *)ws(c, 520, ns * sizeof(Slot));
    if (!o.slot) return ws_fail(c);
    kb_zero<<<1, 64, 0, s>>>(c->flags + F_PSLOT, 1);
    CRV_HIP(c, hipEventRecord(kev[4], s));
    if (nruns) kb_fit<<<fit_blocks, 64, 0, s>>>(p, c->t, g.task_key, roff, o, nruns, c->flags);
    CRV_HIP(c, hipGetLastError());
    CRV_HIP(c, hipEventRecord(kev[5], s));
    CRV_HIP(c, hipMemcpyAsync(hf, c->flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    CRV_HIP(c, hipStreamSynchronize(s));
    if (!hf[F_PSLOT]) break;
    if (attempt >= 5) return set_error(c, CRV_E_OVERFLOW, "batched chain: pulse slots still too few");
    o.pcap *= 2;
  }
  if (opt->pulse_slots <= 0) c->b_pcap = o.pcap;
  CRV_HIP(c, hipEventRecord(ev[2], s));
  // pulse table
  int64_t *npr64 = (int64_t *)ws(c, 541, (nruns + 1) * 8), *poff = (int64_t *)ws(c, 542, (nruns + 2) * 8);
  if (!npr64 || !poff) return ws_fail(c);
  if (nruns) kb_widen<<<nblocks(nruns, 256), 256, 0, s>>>(o.npr, npr64, nruns);
  if ((rc = scan_i64(c, npr64, poff, nruns, 511, s))) return rc;
  int64_t np = 0;
  if ((rc = read_i64(c, poff + nruns, &np, s))) return rc;
  crv_pulses P{};
  const int64_t m = np + 1;
  int64_t **ip[] = {&P.wf, &P.seq, &P.ev};
  for (int i = 0; i < 3; ++i) *ip[i] = (int64_t *)ws(c, 550 + i, m * 8);
  P.chan = (int32_t *)ws(c, 553, m * 4);
  double **dp[] = {&P.pe, &P.peph, &P.time, &P.height, &P.beta, &P.chi2, &P.letime, &P.ped, &P.fit_a, &P.fit_mu,
                   &P.fit_b};
  for (int i = 0; i < 11; ++i) *dp[i] = (double *)ws(c, 554 + i, m * 8);
  uint8_t **bp[] = {&P.failed, &P.zerondf, &P.pedfromdb};
  for (int i = 0; i < 3; ++i) *bp[i] = (uint8_t *)ws(c, 565 + i, m + 8);
  int32_t *pbar = (int32_t *)ws(c, 568, m * 4), *psipm = (int32_t *)ws(c, 569, m * 4);
  bool okp = P.chan && pbar && psipm;
  for (int i = 0; i < 3; ++i) okp = okp && *ip[i] && *bp[i];
  for (int i = 0; i < 11; ++i) okp = okp && *dp[i];
  if (!okp) return ws_fail(c);
  P.capacity = np;
  P.n = np;
  if (nruns) kb_pulses_out<<<nblocks(nruns, 256), 256, 0, s>>>(p, c->t, nruns, g.task_key, poff, o, P, pbar, psipm);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, hipEventRecord(ev[3], s));
  crv_clusters C{};
  C.capacity = -1;
  if ((rc = clusters_stage(c, P.ev, pbar, psipm, P.time, P.pe, P.seq, np, &C, s))) return rc;
  CRV_HIP(c, hipEventRecord(ev[4], s));
  CRV_HIP(c, hipEventSynchronize(ev[4]));
  c->last_pulses = P;
  c->last_clusters = C;
  opt->record_adc = o.rec_adc;
  opt->record_label = o.rec_label;
  opt->record_tick = o.rec_tick;
  opt->n_records = nrec;
  opt->n_waveforms = nruns;
  opt->overlong = hf[F_OVERLONG];
  opt->capped = hf[F_CAPPED];
  opt->pulse_slots_used = o.pcap;
  opt->log_passes = npass;
  if (timing) {
    float ms[4], km[3] = {0.f, 0.f, 0.f}, kc = 0.f, tot = 0.f;
    for (int i = 0; i < 4; ++i) (void)hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
    km[0] = ms_sipm;
    km[1] = ms_wave;
    (void)hipEventElapsedTime(&km[2], kev[4], kev[5]);
    if (np && c->kev_on) (void)hipEventElapsedTime(&kc, kev[6], kev[7]);
    (void)hipEventElapsedTime(&tot, ev[0], ev[4]);
    *timing = crv_timing{};
    timing->charges_ms = ms[0];   // grouping, photon counts, SiPM order
    timing->digis_ms = ms[1];     // the three phases with their sizing and re-runs
    timing->pulses_ms = ms[2];    // pulse table
    timing->clusters_ms = ms[3];
    timing->total_ms = tot;
    timing->k_sipm_ms = km[0];    // kb_sipm (the passes of the last run)
    timing->k_digi_ms = km[1];    // kb_wrec (the passes of the last run)
    timing->k_pulse_ms = km[2];   // kb_fit (last run)
    timing->k_coinc_ms = kc;
    timing->n_steps = n;
    timing->n_hit_counters = T;
    timing->n_photons = hsum[1];
    timing->n_charges = hsum[0];  // fired pixels (the charges, not materialised as a table)
    timing->n_records = nused;    // records with digis (the pool's extent is opts->n_records)
    timing->n_waveforms = nruns;
    timing->n_pulses = np;
    timing->n_clusters = C.n;
    timing->n_members = C.n_mem;
  }
  return CRV_OK;
}


// The batched chain as one fixed-capacity sequence without host synchronisation (crv_graph_*): single pass, pools,
// slots, log and pulse table at the plan's capacities, every count read on the device. Results equal batched_stage's
// whenever no capacity is exceeded; otherwise pl.skip is set and the rest of the call does nothing.
int batched_fixed(Ctx *c, Plan &pl, Arena &ar, hipStream_t s) {
  const crv_params &p = c->p;
  const int W = p.window_samples;
  if (W > 32 * DWORDS || p.samples_zs < 1 || W <= p.samples_zs)
    return set_error(c, CRV_E_ARG, "batched chain: window_samples must be in (samples_zs, 192]");
  if (!c->t.tmpl_f) return set_error(c, CRV_E_STATE, "template not uploaded");
  if (c->t.ntmpl > TMPL_LDS || c->d.nreach > RINGB)
    return set_error(c, CRV_E_ARG, "batched chain: single-PE template too long");
  if (pl.tile % 64 || pl.tile <= 0) return set_error(c, CRV_E_ARG, "batched chain: tile must be a multiple of 64");
  if (pl.grid_s <= 0) pl.grid_s = resident_blocks(c, (const void *)kb_sipm);
  if (pl.grid_r <= 0) pl.grid_r = resident_blocks(c, (const void *)kb_wrec);
  StepGroups g;
  int rc = group_steps_fixed(c, pl, ar, g, s);
  if (rc) return rc;
  const int64_t n = pl.n_cap, L4 = 4 * n, ngrp = (L4 + 63) / 64;
  auto A = [&](size_t b) { return ar.get(b); };
  int32_t *nph = (int32_t *)A(n * 16 + 16);
  int64_t *sipm_nph = (int64_t *)A((L4 + 1) * 8);
  int64_t *gcap = (int64_t *)A((ngrp + 1) * 8), *gsize = (int64_t *)A((ngrp + 1) * 8);
  int64_t *goff = (int64_t *)A((ngrp + 2) * 8);
  uint32_t *skey = (uint32_t *)A(L4 * 4 + 8), *skey2 = (uint32_t *)A(L4 * 4 + 8);
  int32_t *sidx = (int32_t *)A(L4 * 4 + 8), *order = (int32_t *)A(L4 * 4 + 8);
  int32_t *nfired = (int32_t *)A(L4 * 4 + 8);
  int *counters = (int *)A(C_INTS * 4);
  Out o{};
  o.xstride = (int64_t)pl.grid_s * 64;
  o.xcur = (uint16_t *)A((size_t)o.xstride * CMAXB * 2);
  o.xcand = (uint16_t *)A((size_t)o.xstride * CMAXB * 2);
  o.xmult = (uint8_t *)A((size_t)o.xstride * CMAXB);
  o.lastt = (float *)A((size_t)o.xstride * CRV_NPIX * CRV_NPIX * 4);
  o.gS = (float *)A(n * 16 + 16);
  o.gHead = (double *)A(n * 32 + 16);
  o.gTb = (double *)A(n * 32 + 16);
  o.gRJ = (uint64_t *)A(n * 32 + 16);
  o.nrun = (int32_t *)A(L4 * 4 + 8);
  o.flog = (float4 *)A(pl.log_cap * 16 + 16);
  o.tbase = (double *)A(L4 * 8 + 8);
  o.rec_cap = pl.rec_cap;
  o.run_cap = pl.run_cap;
  o.rec_adc = (uint16_t *)A(o.rec_cap * W * 2);
  o.rec_label = (int64_t *)A(o.rec_cap * 8);
  o.rec_tick = (int32_t *)A(o.rec_cap * 4);
  o.run_rec = (int32_t *)A(o.run_cap * 4);
  o.run_i0 = (int32_t *)A(o.run_cap * 4);
  o.run_n = (int32_t *)A(o.run_cap * 4);
  o.run_k = (int32_t *)A(o.run_cap * 4);
  o.run_g = (int64_t *)A(o.run_cap * 8);
  o.run_tdc = (int64_t *)A(o.run_cap * 8);
  o.run_c = (int32_t *)A(o.run_cap * 4);
  o.cd0 = (int32_t *)A(L4 * 4 + 8);
  o.ncls = (int32_t *)A(L4 * 4 + 8);
  o.cdesc_cap = pl.desc_cap;
  o.cdesc = (ClDesc *)A(o.cdesc_cap * sizeof(ClDesc));
  o.ncr = (int32_t *)A(o.cdesc_cap * 4);
  o.crl = (int32_t *)A(o.cdesc_cap * 4);
  int64_t *sums = (int64_t *)A(16), *fired64 = (int64_t *)A(L4 * 8 + 8);
  int64_t *nrun64 = (int64_t *)A((L4 + 1) * 8), *roff = (int64_t *)A((L4 + 2) * 8);
  const int fit_blocks = (int)std::min<int64_t>(nblocks(o.run_cap, 64), 8192);
  o.yscr = (float *)A((size_t)fit_blocks * 64 * W * 4 + 8);
  o.npr = (int32_t *)A(o.run_cap * 4 + 8);
  o.rmap = (int32_t *)A(o.run_cap * 4 + 8);
  o.pcap = pl.pcap;
  o.slot = (Slot *)A((o.run_cap * o.pcap + 8) * sizeof(Slot));
  int64_t *npr64 = (int64_t *)A((o.run_cap + 1) * 8), *poff = (int64_t *)A((o.run_cap + 2) * 8);
  int32_t *pbar = (int32_t *)A(pl.np_cap * 4 + 8), *psipm = (int32_t *)A(pl.np_cap * 4 + 8);
  void *ptrs[] = {nph, sipm_nph, gcap, gsize, goff, skey, skey2, sidx, order, nfired, counters, o.xcur, o.xcand,
                  o.xmult, o.lastt, o.gS, o.gHead, o.gTb, o.gRJ, o.nrun, o.flog, o.tbase, o.rec_adc, o.rec_label,
                  o.rec_tick, o.run_rec, o.run_i0, o.run_n, o.run_k, o.run_g, o.run_tdc, sums, fired64, nrun64, roff,
                  o.yscr, o.npr, o.rmap, o.slot, npr64, poff, pbar, psipm, o.run_c, o.cd0, o.ncls, o.cdesc, o.ncr,
                  o.crl};
  for (void *q : ptrs)
    if (!q) return set_error(c, CRV_E_HIP, "graph: buffer allocation failed" + ar.note());
  o.rec_count = counters + C_REC;
  o.run_count = counters + C_RUN;
  o.cdesc_count = counters + C_DESC;
  o.rec_used = counters + C_USED;
  o.goff = goff;
  o.gcap = gcap;
  o.log0 = 0;
  kb_zero<<<1, 64, 0, s>>>(counters, C_INTS);
  kb_zero<<<1, 64, 0, s>>>(pl.skip, 2);
  const BStep st{g.edep, g.z, g.t};
  kb_count<<<nblocks(L4, 256), 256, 0, s>>>(p, c->t, c->d, st, pl.T_dev, g.task_key, g.task_off, g.perm, g.evfirst,
                                            nph, sipm_nph, skey, sidx, pl.sort_tasks, c->flags);
  const int sbits = sort_key_bits(pl.sort_tasks);
  const uint32_t pad_key = sbits >= 32 ? 0xFFFFFFFFu : (1u << sbits) - 1u;
  kb_pad_sipms<<<nblocks(L4, 256), 256, 0, s>>>(L4, pl.T_dev, pad_key, skey, sidx, sipm_nph, nfired, o.nrun);
  CRV_HIP(c, hipGetLastError());
  const bool sorted = pl.sort_tasks != 0;
  if (sorted) {
    size_t tb = 0;
    CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, skey, skey2, sidx, order, (int)L4, 0, sbits, s));
    void *tmp = A(tb);
    if (!tmp) return set_error(c, CRV_E_HIP, "graph: sort buffer allocation failed" + ar.note());
    CRV_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, skey, skey2, sidx, order, (int)L4, 0, sbits, s));
  }
  const int32_t *qorder = sorted ? order : nullptr;
  kb_gcaps<<<nblocks(ngrp * 64, 256), 256, 0, s>>>(sipm_nph, qorder, L4, gcap, gsize, pl.log_factor, pl.T_dev);
  CRV_HIP(c, hipGetLastError());
  if ((rc = scan_fixed(c, ar, gsize, goff, ngrp, s))) return rc;
  kb_check_log<<<1, 1, 0, s>>>(goff + ngrp, pl.log_cap, pl.skip);
  auto mark = [&](int k) -> hipError_t { return pl.kev[k] ? hipEventRecord(pl.kev[k], s) : hipSuccess; };
  CRV_HIP(c, mark(0));
  kb_sipm<<<pl.grid_s, 64, 0, s>>>(p, c->t, c->d, st, g.task_key, g.task_off, g.perm, g.evfirst, nph, o,
                                   Queue{counters, pl.tile, L4, qorder, 0, pl.T_dev, pl.skip}, nfired, c->flags);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, mark(1));
  kb_dmark<<<1, 1, 0, s>>>(o.cdesc_count, counters + C_DRANGE + 1);  // descriptors [0, counters[C_DRANGE + 1])
  CRV_HIP(c, mark(2));
  kb_wrec<<<pl.grid_r, 64, 0, s>>>(p, c->t, c->d, g.task_key, o, counters + C_DRANGE, pl.skip, c->flags);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, mark(3));
  kb_check_wave<<<1, 1, 0, s>>>(c->flags, counters, o.rec_cap, o.run_cap, o.cdesc_cap, pl.skip);
  // photons drawn and pixels fired
  kb_widen<<<nblocks(L4, 256), 256, 0, s>>>(nfired, fired64, L4);
  {
    size_t tb = 0;
    CRV_HIP(c, hipcub::DeviceReduce::Sum(nullptr, tb, fired64, sums, (int)L4, s));
    void *tmp = A(tb);
    if (!tmp) return set_error(c, CRV_E_HIP, "graph: reduce buffer allocation failed" + ar.note());
    CRV_HIP(c, hipcub::DeviceReduce::Sum(tmp, tb, fired64, sums, (int)L4, s));
    CRV_HIP(c, hipcub::DeviceReduce::Sum(tmp, tb, sipm_nph, sums + 1, (int)L4, s));
  }
  kb_run_order<<<nblocks(L4, 256), 256, 0, s>>>(L4, pl.T_dev, o, pl.skip);
  kb_run_rank<<<nblocks(o.run_cap, 256), 256, 0, s>>>(o, o.run_cap, o.run_count, pl.skip);
  kb_widen<<<nblocks(L4, 256), 256, 0, s>>>(o.nrun, nrun64, L4);
  if ((rc = scan_fixed(c, ar, nrun64, roff, L4, s))) return rc;
  kb_zero32<<<nblocks(o.run_cap + 1, 256), 256, 0, s>>>(o.npr, o.run_cap + 1);
  kb_zero<<<1, 64, 0, s>>>(c->flags + F_PSLOT, 1);
  CRV_HIP(c, mark(4));
  kb_fit<<<fit_blocks, 64, 0, s>>>(p, c->t, g.task_key, roff, o, o.run_cap, c->flags, o.run_count, pl.skip);
  CRV_HIP(c, hipGetLastError());
  CRV_HIP(c, mark(5));
  kb_widen<<<nblocks(o.run_cap, 256), 256, 0, s>>>(o.npr, npr64, o.run_cap);
  if ((rc = scan_fixed(c, ar, npr64, poff, o.run_cap, s))) return rc;
  kb_check_fit<<<1, 1, 0, s>>>(c->flags, poff + o.run_cap, pl.np_cap, pl.skip);
  crv_pulses P = pl.P;
  P.capacity = pl.np_cap;
  kb_pulses_out<<<nblocks(o.run_cap, 256), 256, 0, s>>>(p, c->t, o.run_cap, g.task_key, poff, o, P, pbar, psipm,
                                                         o.run_count, pl.skip);
  CRV_HIP(c, hipGetLastError());
  kb_pad_pulses<<<nblocks(pl.np_cap, 256), 256, 0, s>>>(pl.np_cap, poff + o.run_cap, pl.skip, P.ev);
  kb_sizes<<<1, 1, 0, s>>>(poff + o.run_cap, sums, counters, pl.skip, pl.sizes);
  CRV_HIP(c, hipGetLastError());
  return clusters_fixed(c, pl, ar, P.ev, pbar, psipm, P.time, P.pe, P.seq, s);
}

}  // namespace crv
