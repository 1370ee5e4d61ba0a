// Internal declarations of libcrvhip: device tables, the context, workspace and the stage launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/crvhip.h"
#include "crv_math.h"

#define CRV_CHAN_MULT 32768
#define CRV_NPIX 41
#define CRV_DEAD_WORDS 53   // ceil(41 * 41 / 32)

namespace crv {

// Read-only device tables uploaded once per context (conditions, template, photon map, finder geometry).
// what the coincidence finders read of a bar, with its sector's fields, in one 64-byte record (one request per hit
// instead of a chain of dependent loads bar -> sector -> sector fields)
struct BarInfo {
  double x, y;      // bar centre (the finder's x and y)
  double ps0, ps1;  // its sector's side-0 and side-1 positions
  double imd;       // its sector's initial_max_distance
  int32_t sec;      // sector
  int32_t type;     // sector type when the sector has cuts, else INT32_MIN (matches no type)
  int32_t layer, cit;
  int32_t pad_[2];
};

struct Tables {
  int32_t n_channels = 0;
  const int32_t *status = nullptr, *roc = nullptr, *feb = nullptr;
  const double *dev = nullptr, *pedestal = nullptr, *area = nullptr, *height = nullptr, *toff = nullptr;
  const double *tmpl = nullptr;
  const float *tmpl_f = nullptr;  // FP32 copy
  int32_t ntmpl = 0;
  double vmax = 0.0;
  const double *cdf = nullptr;
  int32_t ncdf = 0, map_ny = 0;
  int32_t map_uniform = 0;  // every map bin is 1: the CDF lookup is floor(x)
  double cdf_total = 0.0;
  const uint32_t *dead = nullptr;  // CRV_DEAD_WORDS words, bit pix = x * 41 + y
  // the dead mask's nonzero words when they lie within 8 consecutive words (dead_w0 >= 0; else -1): kernels test a
  // pixel against these in registers instead of loading the mask
  int32_t dead_w0 = -1;
  uint32_t dead_win[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // finder geometry
  int32_t nbars = 0, nsectors = 0, ntypes = 0;
  int32_t ncit = 0;  // 1 + the largest counter_in_type
  const int32_t *sector = nullptr, *layer = nullptr, *cit = nullptr;
  const double *bx = nullptr, *by = nullptr, *bpos = nullptr;
  const int32_t *stype = nullptr, *ldir = nullptr;
  const double *ps0 = nullptr, *ps1 = nullptr;
  const uint8_t *sside0 = nullptr, *sside1 = nullptr, *hascuts = nullptr;
  const double *cuts = nullptr;  // [nsectors * 10] SectorCuts fields in declaration order
  const BarInfo *binfo = nullptr;  // [nbars]
  // per sector type (sorted type ids): p0, p1, half, max_dt, side0, side1
  const int32_t *type_id = nullptr;
  const double *tp0 = nullptr, *tp1 = nullptr, *thalf = nullptr, *tmaxdt = nullptr;
  const uint8_t *tside0 = nullptr, *tside1 = nullptr;
};

// A v6 lookup table on the device (crv_lookup_upload): constants, edges, Cerenkov maps, per-bin probabilities and
// cumulative delay weights (int32 [bins * dmax]).
struct LkTab {
  double ht, hw, hl, fs, hrx, hry, fr, cr, inv_v, att;
  const double *edges[7];
  int32_t ne[7], nb[7];
  const double *ck[2];
  int32_t nck[2];
  const float *prob[3];
  const int32_t *cum[3], *dsum[3];
  int32_t dmax[3];
};
struct LkBars {
  int32_t nbars = 0;
  const int32_t *table = nullptr, *refl = nullptr;
  const double *yield = nullptr, *speed = nullptr, *atten = nullptr, *center = nullptr;
  const int32_t *l2w = nullptr;
};

// Derived scalars of the parameters (computed on the host with the same portable functions as the device).
struct Derived {
  double yscale;  // pe_per_mev / A(ref_distance)
  double q_ct;    // per-neighbour cross-talk candidate probability
  int32_t nreach;
};

struct Ctx {
  int device = 0;
  crv_params p{};
  bool have_params = false;
  Derived d{};
  Tables t;
  std::vector<void *> owned;  // device allocations of the tables
  std::vector<void *> ws;     // workspace slots
  std::vector<size_t> ws_size;
  int *flags = nullptr;       // device error / diagnostic counters
  std::string err;
  hipEvent_t ev[8];
  crv_pulses last_pulses{};      // tables of the last chained run (workspace)
  crv_clusters last_clusters{};
  crv_digis last_digis{};
  hipEvent_t kev[8];             // start / end of the main kernel of each stage (sipm, digi, pulse, coinc)
  hipStream_t aux[2] = {nullptr, nullptr};  // side streams of the finder's size classes (default priority)
  std::vector<hipEvent_t> pev;               // per-pass events of the pipelined batched chain (grown on demand)
  hipEvent_t fork[3];                       // fork / join events of the side streams
  bool kev_on = false;
  int64_t b_rec_cap = 0, b_run_cap = 0, b_desc_cap = 0;  // pool sizes of the batched chain (grown on demand)
  int b_log_factor = 1;  // fired-pixel log capacity factor of the batched chain, doubled on overflow and kept
  int b_pcap = 0;        // pulse slots per waveform the batched chain grew to (kept for later calls)
  int64_t b_last_log = 0;  // 16-byte chunks of the whole fired-pixel log of the last batched call
  int64_t b_last_desc = 0; // kept charge clusters of the last batched call
  std::vector<int> ws_spare;  // slots the running call no longer reads: freed to make room when an allocation fails
  int ws_fail_slot = -1;                  // the workspace request that failed last
  size_t ws_fail_bytes = 0;
  uint8_t *scratch = nullptr;  // this device's fixed scratch array of the allocation-free pulse entry point
  // lookup-table optics (crv_lookup_upload)
  std::vector<void *> lk_owned;
  const LkTab *lk_tabs = nullptr;
  int32_t lk_ntab = 0;
  LkBars lk_bars{};
};

// Workspace slots: each stage takes fixed slots below WS_STAGE_SLOTS (charges 0-25, digis 30-55, pulses 60-81,
// reco-waveform grouping 130-138, finder 90-191, reusing 130-138 whose grouping nothing reads by then; tables of the
// chained FP64 run 300-347; batched chain 500-583). A scan on slot k keeps its temporary storage in WS_SCAN_TMP + k
// and its widened input in WS_SCAN_WIDE + k, so a scan never lands on a table that a stage still holds.
constexpr int WS_STAGE_SLOTS = 2000, WS_SCAN_TMP = 2000, WS_SCAN_WIDE = 4000;

// workspace slot of at least `bytes` bytes (grown, never shrunk; when the device has no room, the slots listed in
// ws_spare are freed once, after the device has drained, and the allocation is tried again)
void *ws(Ctx *c, int slot, size_t bytes);

// device flag words
enum Flag { F_UNSORTED = 0, F_COUNTER = 1, F_CAPPED = 2, F_OVERFLOW = 3, F_CASCADE = 4, F_OVERLONG = 5, F_BADCHAN = 6,
            F_BADINPUT = 8, F_NFLAGS = 10 };

int set_error(Ctx *c, int code, const std::string &msg);
// CRV_E_HIP with the workspace request that could not be allocated
int ws_fail(Ctx *c);
#define CRV_HIP(c, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return set_error((c), CRV_E_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// ---- stage launchers (host functions; all asynchronous on `s`) ----
struct StepGroups {  // hit counters: steps grouped by (event, counter)
  int64_t n_steps = 0, T = 0;
  const uint32_t *ev = nullptr, *counter = nullptr;
  const double *edep = nullptr, *z = nullptr, *t = nullptr;
  uint64_t *task_key = nullptr;  // (event << 13) | counter
  int32_t *task_off = nullptr;   // [T + 1]
  int32_t *perm = nullptr;       // step rows in task order
  int32_t *evfirst = nullptr;    // first row of each step's event
};
int group_steps(Ctx *c, const crv_steps *st, StepGroups &g, hipStream_t s);
int charges_stage(Ctx *c, StepGroups &g, crv_charges *out, hipStream_t s);
// the photoelectrons of the steps as a list (crv_photons_run), optionally shifted per event and folded by a period
int photons_stage(Ctx *c, StepGroups &g, const double *shift, double period, crv_photons *out, hipStream_t s);
// lookup-table optics (k_lookup.hip)
int lookup_upload(Ctx *c, const crv_lookup_table *tabs, int32_t ntab, const crv_lookup_bars *bars);
void lookup_free(Ctx *c);
int lookup_photons_stage(Ctx *c, const crv_lookup_steps *st, crv_photons *out, hipStream_t s);
// photon list + dark noise -> charges with cross-talk and trap afterpulses (crv_sipm_run)
int sipm_list_stage(Ctx *c, const crv_photon_list *ph, const crv_sipm_noise *nz, crv_charges *out, hipStream_t s);
int digis_stage(Ctx *c, const int64_t *label, const double *time, const double *charge, int64_t n, crv_digis *out,
                hipStream_t s);
int waveforms_stage(Ctx *c, const int64_t *label, const int64_t *starttdc, const int16_t *adc, int64_t n,
                    crv_waveforms *out, hipStream_t s);
int pulses_stage(Ctx *c, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                 const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                 const double *height, const double *toff, crv_pulses *out, hipStream_t s);
// the same on the device's fixed scratch array (no allocation; sizes bounded, see k_pulse.hip); noalloc_setup looks
// up the array's address on the current device (once per device) and keeps it in the context
int pulses_stage_noalloc(Ctx *c, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                         const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                         const double *height, const double *toff, crv_pulses *out, hipStream_t s);
int noalloc_setup(Ctx *c);
int clusters_stage(Ctx *c, const int64_t *ev, const int32_t *bar, const int32_t *sipm, const double *time,
                   const double *pe, const int64_t *seq, int64_t n, crv_clusters *out, hipStream_t s);
// the FP32 batched chain; pulses and clusters end in last_pulses / last_clusters
int batched_stage(Ctx *c, const crv_steps *steps, crv_batch_opts *opt, crv_timing *timing, hipStream_t s);

// ---- the captured (fixed-capacity) batched chain of crv_graph_* ----
// Device buffers of a fixed-capacity call: the first (eager) run allocates them in order; while the same sequence is
// captured into a graph, the same requests receive the same buffers again, so capture allocates nothing.
struct Arena {
  std::vector<std::pair<void *, size_t>> blocks;
  size_t next = 0;
  bool replay = false;
  size_t failed = 0;  // bytes of the last request that could not be met
  void *get(size_t bytes);
  void release();
  std::string note() const;  // " (request, held, device free)" for allocation errors
};

// Capacities of a fixed-capacity batched call and its device-side counts. Every size a kernel or a library call needs
// is a capacity known at capture; the counts of a call live in device memory, the kernels read them there, and
// whatever exceeds a capacity sets `skip`, which empties the rest of the call (the host runs it uncaptured instead).
struct Plan {
  int64_t n_cap = 0, log_cap = 0, rec_cap = 0, run_cap = 0, np_cap = 0, desc_cap = 0;
  int pcap = 4, log_factor = 1, tile = 64, sort_tasks = 1, grid_s = 0, grid_r = 0;
  // the graph's inputs: steps at capacity and their count
  uint32_t *in_event = nullptr, *in_counter = nullptr;
  double *in_edep = nullptr, *in_z = nullptr, *in_t = nullptr;
  int64_t *n_dev = nullptr;
  int32_t *T_dev = nullptr;  // hit counters of the call
  int *skip = nullptr;       // [0] skip (a capacity was exceeded), [1] which: 1 log, 2 pools, 3 pulse slots, 4 pulses
  int64_t *sizes = nullptr;  // [0] pulses, [1] clusters, [2] members, [3] photons, [4] fired pixels, [5] records,
                             // [6] runs
  crv_pulses P{};            // outputs at capacity np_cap (clusters and members are bounded by the pulses)
  crv_clusters C{};
  hipEvent_t kev[6] = {};    // start / end of kb_sipm, kb_wave and kb_fit (recorded when set)
};
int group_steps_fixed(Ctx *c, Plan &pl, Arena &ar, StepGroups &g, hipStream_t s);
int batched_fixed(Ctx *c, Plan &pl, Arena &ar, hipStream_t s);
int clusters_fixed(Ctx *c, Plan &pl, Arena &ar, const int64_t *ev, const int32_t *bar, const int32_t *sipm,
                   const double *time, const double *pe, const int64_t *seq, hipStream_t s);
// device-side exclusive scan over a fixed item count (n + 1 outputs, first 0), temporaries from the arena
int scan_fixed(Ctx *c, Arena &ar, const int64_t *in, int64_t *out, int64_t n, hipStream_t s);
int widen_scan_fixed(Ctx *c, Arena &ar, const int32_t *in, int64_t *out, int64_t n, hipStream_t s);

// device-side exclusive scan of int64 counts (n + 1 outputs, last = total) with hipcub
int scan_i64(Ctx *c, const int64_t *in, int64_t *out, int64_t n, int slot, hipStream_t s);
int scan_i32_to_i64(Ctx *c, const int32_t *in, int64_t *out, int64_t n, int slot, hipStream_t s);
int read_i64(Ctx *c, const int64_t *dev, int64_t *host, hipStream_t s);

inline unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// HIP events around the main kernel `k` (0 sipm, 1 digi, 2 pulse, 3 coinc) when the chained run asks for them
inline void kmark(Ctx *c, int k, int end, hipStream_t s) {
  if (c->kev_on) (void)hipEventRecord(c->kev[2 * k + end], s);
}

}  // namespace crv
