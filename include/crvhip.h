/* libcrvhip: the CRV response chain of crvfastsim (CrvSteps -> photoelectrons -> SiPM charges -> digitised
 * waveforms -> pulses -> coincidence clusters) as HIP kernels for MI355X (gfx950), behind a plain C ABI.
 *
 * The reference has no FFI for this path: its operator API is the Python stage functions of crvfastsim/chain.py
 * (charges :151, digis :234, reco_waveforms :273, clusters :296). Each entry point below replaces one of them (or a
 * fused run of them), with the reference's tables as device arrays:
 *   - every call returns 0 or a negative CRV_E* status; crv_last_error() gives the message;
 *   - the caller owns every input and output buffer (device pointers, e.g. torch tensors); the library never frees
 *     caller memory and keeps only its own workspace;
 *   - outputs have a caller-given capacity; when a table would not fit, the call writes nothing, stores the size
 *     it needs in the `n` field and returns CRV_E_CAPACITY (never a truncated table: the reference's
 *     response/contract.py:24-29 require_complete rule);
 *   - one context per host thread; calls are ordered on the given stream and return after it has finished.
 */
#ifndef CRVHIP_H
#define CRVHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRV_OK 0
#define CRV_E_ARG (-1)
#define CRV_E_HIP (-2)
#define CRV_E_CAPACITY (-3)
#define CRV_E_STATE (-4)
#define CRV_E_OVERFLOW (-5)

typedef struct crv_ctx crv_ctx;
typedef void *crv_stream; /* a hipStream_t; NULL is the null stream */

/* Physics and reconstruction settings: the fields of the reference's SiPMConfig (response/sipm.py:26-47),
 * WaveformConfig (reco/waveform.py:29-44), RecoConfig (reco/pulsefit.py:26-46), FinderConfig (reco/coinc.py:52-61)
 * and ChainConfig (chain.py:61-84), plus the closed-form optics and the Philox seed. */
typedef struct crv_params {
  uint64_t seed;
  uint32_t evkey_hi; /* event key >> 32, shared by the events of a call */
  int32_t precision; /* 64: plain FP64 stages; 32: FP32 batched chain */
  /* optics */
  double pe_per_mev, ref_distance, w_short, l_short, w_long, l_long, half_length, inv_speed, tau;
  double photon_window_hi; /* photons after it are dropped (ChainConfig.window[1]) */
  /* SiPM */
  int32_t n_pixels_x, n_pixels_y, pixel_lrint, _pad0;
  double overvoltage, time_constant, capacitance, p1, p2, cross_talk;
  /* waveforms */
  double single_pe_reference_charge, single_pe_max_time, single_pe_precision, min_voltage, noise;
  int32_t noise_on, samples_zs, window_samples, feb_key_roc;
  double adc_conversion, window_lo, window_hi, feb_phase_width;
  int32_t pedestal, min_adc, max_adc, _pad1;
  /* pulse reco */
  double min_adc_difference, default_beta, min_beta, max_beta, max_time_difference, min_pulse_height_ratio;
  double max_pulse_height_ratio, le_time_factor, pedestal_undershoot_threshold, time_offset_scale, limit_tolerance;
  int32_t samples_before, samples_after, use_time_offset_db, ignore_channels, lm_iterations, _pad2;
  /* coincidence finder */
  double fiber_signal_speed, max_time_interval, time_offset;
  int32_t compensate_mask, ignore_secondary_pulses, big_cluster_threshold, add_opposite_side_pulses;
} crv_params;

/* Context on a device (the library's HIP device id). */
int crv_ctx_create(int device, crv_ctx **out);
int crv_ctx_destroy(crv_ctx *ctx);
int crv_set_params(crv_ctx *ctx, const crv_params *p);
/* Message of the last failed call on this context (or of the last failed crv_ctx_create when ctx is NULL). */
int crv_last_error(crv_ctx *ctx, char *buf, int64_t size);

/* Per-channel conditions (host arrays): reference chain.py:29-52 Conditions and readout.py:58-72 effective status;
 * dev = the photon-yield factor (deviation * variation scale + 1) * yield scale of response/photons.py:348. */
int crv_upload_conditions(crv_ctx *ctx, int32_t n_channels, const int32_t *status, const int32_t *roc,
                          const int32_t *feb, const double *dev, const double *pedestal, const double *area,
                          const double *height, const double *toff);
/* Resampled single-PE template in volts on the single_pe_precision grid (reco/waveform.py:46-62). */
int crv_upload_template(crv_ctx *ctx, const double *tmpl, int32_t n);
/* Photon map [nx * ny] (response/sipm.py:49-91) and the dead-pixel mask [41 * 41] (SiPMConfig.inactive). */
int crv_upload_photon_map(crv_ctx *ctx, const double *pmap, int32_t nx, int32_t ny, const uint8_t *dead41);
/* Coincidence-finder geometry (reco/coinc.py:64-122 SectorGeom and SectorCuts): per bar sector, layer,
 * counter_in_type, x, y, pos[3]; per sector type, length direction, readout-end positions, SiPM sides and the
 * ten SectorCuts values (has_cuts = 0 for sectors without cuts). */
int crv_upload_geometry(crv_ctx *ctx, int32_t nbars, const int32_t *sector, const int32_t *layer,
                        const int32_t *counter_in_type, const double *x, const double *y, const double *pos,
                        int32_t nsectors, const int32_t *sector_type, const int32_t *length_dir,
                        const double *pos_side0, const double *pos_side1, const uint8_t *sipms_side0,
                        const uint8_t *sipms_side1, const uint8_t *has_cuts, const double *cuts);

/* CrvSteps: SoA device arrays sorted by event (32 bytes per step). */
typedef struct crv_steps {
  const uint32_t *event;
  const uint32_t *counter;
  const double *edep;
  const double *z;
  const double *t;
  int64_t n;
} crv_steps;

/* SiPM charges (reference chain.charges return dict), sorted by (label, time). */
typedef struct crv_charges {
  int64_t *label; /* event key * 32768 + channel */
  double *time, *charge, *pe;
  int64_t *photon; /* step row */
  int64_t capacity, n;
  int64_t n_photons, capped; /* photoelectrons drawn; Poisson draws that hit the rejection bound */
} crv_charges;

/* Replaces chain.charges (chain.py:151-213, numpy branch): photons (response/photons.py) + SiPM charges
 * (response/sipm.py). FP64, one launch per stage. */
int crv_charges_run(crv_ctx *ctx, const crv_steps *steps, crv_charges *out, crv_stream stream);

/* Photon arrivals at the SiPMs (device arrays): SiPM label (event key * 32768 + channel), time, and an id carried to
 * the charges. A photon's draws are keyed by (label, j, rank): the step ordinal and rank crv_photons_run gives it,
 * or, when j is NULL, its position in the list and rank 0. */
typedef struct crv_photon_list {
  const int64_t *label;
  const double *time;
  const int64_t *photon;
  const uint32_t *j, *rank; /* optional (NULL) */
  int64_t n;
} crv_photon_list;

/* Photon arrivals of CrvSteps: the photon half of crv_charges_run (reference response/photons.py:102-224 with the
 * closed-form optics, status mask chain.py:199-204). Every photoelectron of a live channel, in (label, step ordinal,
 * rank) order, with its step row and draw identity; with `shift` (device array indexed by the steps' event), times
 * become fmod(t + shift[event], period) folded into [0, period) (reference onspill.py:117-118 and :145-148). */
typedef struct crv_photons {
  int64_t *label;
  double *time;
  int64_t *photon; /* step row */
  uint32_t *j, *rank;
  int64_t capacity, n;
  int64_t capped; /* Poisson draws that hit the rejection bound */
} crv_photons;

int crv_photons_run(crv_ctx *ctx, const crv_steps *steps, const double *shift, double period, crv_photons *out,
                    crv_stream stream);

/* Stock lookup-table optics (reference response/lookup.py:65-96 LookupTable, read from the v6 file by
 * crvhip.lookup.load): one table's geometry constants, Cerenkov maps, bin edges and, for each of its three tables
 * (scintillation in the scintillator, Cerenkov in the scintillator, Cerenkov in the fibre), the per-bin arrival
 * probability and the delay weights padded to dmax per bin; inv_speed / atten_length as
 * response/photons.py:230-254 table_propagation measures them. Host arrays; the library keeps device copies. */
typedef struct crv_lookup_table {
  double half_thickness, half_width, half_length, fiber_separation, hole_radius_x, hole_radius_y, fiber_radius,
      corner_radius;
  double inv_speed, atten_length;
  const double *edges[7]; /* x, y, z, beta, theta, phi, r */
  int32_t n_edges[7];
  const double *ck[2];    /* [n, 2] (beta, photons per mm): scintillator, fibre */
  int32_t n_ck[2];
  const float *prob[3];   /* [bins] */
  const uint8_t *delays[3]; /* [bins * dmax] */
  int32_t dmax[3];
} crv_lookup_table;

/* Per bar (response/photons.py:305-356 iter_rows inputs): the table index of its sector (-1: not simulated), the
 * sector's reflector (0: both ends read; otherwise SiPMs 0 and 2 only, reflector < 0 writing SiPM s to channel
 * s + 1), its yield (photons per MeV), the propagation ratios of ResponseConfig.speed_ratio / atten_ratio, its
 * centre (mm) and local_to_world (the world axes of thickness, width, length). */
typedef struct crv_lookup_bars {
  int32_t nbars;
  const int32_t *table, *reflector;
  const double *yield, *speed_ratio, *atten_ratio;
  const double *center;        /* [nbars * 3] */
  const int32_t *local_to_world; /* [nbars * 3] */
} crv_lookup_bars;

/* Upload 1-16 tables and the bar arrays (replaces any earlier upload on the context). */
int crv_lookup_upload(crv_ctx *ctx, const crv_lookup_table *tables, int32_t n_tables, const crv_lookup_bars *bars);

/* CrvSteps for the lookup optics: SoA device arrays sorted by event (reference CrvStep columns, README.md:30-40). */
typedef struct crv_lookup_steps {
  const uint32_t *event;        /* event key low 32 bits (the upper 32 in crv_params.evkey_hi) */
  const int32_t *bar;
  const double *t0, *t1;
  const double *pos0, *pos1;    /* [n * 3] world (x, y, z) of the step's start and end, mm */
  const double *edep, *len;     /* MeV, mm */
  const double *p0, *p1;        /* momentum magnitude at the start and end, MeV */
  const double *mass, *charge;  /* particle mass (MeV) and charge (e) of the step's pdg (photons.py:65-77) */
  int64_t n;
} crv_lookup_steps;

/* Replaces response/photons.py:359-371 make_photons (the substep loop of :163-224) with the status mask of
 * chain.py:199-204: the photoelectrons of the steps in (step, SiPM, rank) order, a row's photons ranked substep by
 * substep, then by type, then by arrival; photon = the step's row in the call, (j, rank) its draw identity for
 * crv_sipm_run. FP64; CRV_E_CAPACITY with out->n set when the output is too small. */
int crv_lookup_photons_run(crv_ctx *ctx, const crv_lookup_steps *steps, crv_photons *out, crv_stream stream);

/* Dark noise and trap afterpulses (reference response/sipm.py:26-47 SiPMConfig thermal_rate, trap*_prob and
 * trap*_lifetime; simulate's t_start, t_end and thermal_chans, sipm.py:102-121). */
typedef struct crv_sipm_noise {
  double thermal_rate;                  /* dark discharges per ns at full overvoltage (0: none) */
  double trap0_prob, trap1_prob;        /* per firing (0: none) */
  double trap0_lifetime, trap1_lifetime;/* ns */
  double t_start, t_end;                /* dark noise in [t_start, t_end); photons and afterpulses after t_end dropped */
  const int64_t *thermal_labels;        /* device array of distinct SiPM labels that get dark noise */
  int64_t n_thermal;
} crv_sipm_noise;

/* Replaces response/sipm.py:102-198 simulate (with thermal_chans, t_start, t_end): photon arrivals plus dark noise ->
 * SiPM charges with cross-talk and trap afterpulses, sorted by (label, time); photon -1 marks dark discharges and
 * their descendants. FP64. out->n_photons counts the photons kept and the dark discharges. */
int crv_sipm_run(crv_ctx *ctx, const crv_photon_list *photons, const crv_sipm_noise *noise, crv_charges *out,
                 crv_stream stream);

/* Zero-suppressed digis (reference chain.digis return dict) and the 128-sample ADC record of every charge cluster
 * that has a digi. */
typedef struct crv_digis {
  int64_t *label;
  double *start, *tdc0;
  int64_t *starttdc;
  int16_t *adc; /* [capacity * samples_zs] */
  int32_t *record; /* record row of each digi */
  int32_t *sample; /* first sample of each digi in its record */
  int64_t capacity, n;
  uint16_t *record_adc; /* [record_capacity * window_samples] */
  int64_t *record_label;
  int32_t *record_tick; /* window start = tick * 12.5 + tdc0 */
  int64_t record_capacity, n_records;
  int64_t overlong; /* clusters longer than the window */
} crv_digis;

/* Replaces chain.digis (chain.py:234-270, numpy branch: tdc0_per_feb + reco/waveform.py synthesize). Input
 * charges grouped by label and time-ordered within it (as crv_charges_run writes them). */
int crv_digis_run(crv_ctx *ctx, const int64_t *label, const double *time, const double *charge, int64_t n,
                  crv_digis *out, crv_stream stream);

/* Reco waveforms (reference chain.reco_waveforms, chain.py:273-286 + reco/pulsefit.py:319-345): continuing
 * digis of one channel concatenated, sorted by (event, channel, start). Samples are flat, waveform w owning
 * samples [off[w], off[w+1]). */
typedef struct crv_waveforms {
  int64_t *ev;
  int32_t *chan;
  int64_t *start; /* start TDC */
  int32_t *nsamp;
  int64_t *off;   /* [capacity + 1] */
  int16_t *adc;   /* [sample_capacity] */
  int64_t capacity, n, sample_capacity, n_samples;
} crv_waveforms;

int crv_reco_waveforms_run(crv_ctx *ctx, const int64_t *label, const int64_t *starttdc, const int16_t *adc,
                           int64_t n, crv_waveforms *out, crv_stream stream);

/* Reco pulses (reference reco/pulsefit.py:236-316 reconstruct_full, then filter_channels and the DB time offset
 * of chain.py:317-323), sorted by (waveform, sequence). */
typedef struct crv_pulses {
  int64_t *wf, *seq;
  int64_t *ev;   /* event key of the pulse's waveform (optional, NULL) */
  int32_t *chan; /* channel of the pulse's waveform (optional, NULL) */
  double *pe, *peph, *time, *height, *beta, *chi2, *letime, *ped, *fit_a, *fit_mu, *fit_b;
  uint8_t *failed, *zerondf, *pedfromdb;
  int64_t capacity, n;
} crv_pulses;

/* Waveforms as flat samples: ev, chan, start TDC, nsamp, off per waveform; calibration per waveform. */
int crv_pulses_run(crv_ctx *ctx, const int64_t *ev, const int32_t *chan, const int64_t *start, const int32_t *nsamp,
                   const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped, const double *area,
                   const double *height, const double *toff, crv_pulses *out, crv_stream stream);

/* crv_pulses_run without any device allocation: no counterpart in the reference. For callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the caller passes inputs and
 * outputs (capacity >= 0) as for crv_pulses_run, same results; all scratch lives in one fixed device array per GPU
 * (about 46 MiB, held from the first context on that GPU until the process ends). Limits per call: 131,072
 * waveforms and 524,288 samples (CRV_E_ARG beyond). Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (two streams, two host threads) race; the call returns after its stream has finished. */
int crv_pulses_run_noalloc(crv_ctx *ctx, const int64_t *ev, const int32_t *chan, const int64_t *start,
                           const int32_t *nsamp, const int64_t *off, const int16_t *adc, int64_t n_wf, const double *ped,
                           const double *area, const double *height, const double *toff, crv_pulses *out,
                           crv_stream stream);

/* Coincidence clusters (reference chain.clusters detail output, chain.py:325-362 + reco/coinc.py find_event). */
typedef struct crv_clusters {
  int64_t *cl_event;
  int32_t *cl_type, *cl_poscalc, *cl_sidehits; /* sidehits [capacity * 2] */
  double *cl_pe, *cl_start, *cl_end, *cl_slope, *cl_avgtime, *cl_pos, *cl_sidepe, *cl_sidetime; /* pos [cap*3] */
  int64_t *cl_mem_off; /* [capacity + 1] */
  int64_t capacity, n;
  int64_t *mem_idx; /* pulse index within its event */
  int64_t mem_capacity, n_mem;
} crv_clusters;

/* Pulses of several events sorted by (event, waveform, sequence): ev, bar, sipm, time, pe, seq per pulse. */
int crv_clusters_run(crv_ctx *ctx, const int64_t *ev, const int32_t *bar, const int32_t *sipm, const double *time,
                     const double *pe, const int64_t *seq, int64_t n, crv_clusters *out, crv_stream stream);

/* Per-stage and per-kernel timings of the last chain call, in milliseconds (HIP events on the chain's stream), and
 * the table sizes of the call. */
typedef struct crv_timing {
  double charges_ms, digis_ms, waveforms_ms, pulses_ms, clusters_ms, total_ms;
  double k_sipm_ms, k_digi_ms, k_pulse_ms, k_coinc_ms; /* the main kernel of each stage */
  int64_t n_steps, n_hit_counters, n_photons, n_charges, n_digis, n_records, n_waveforms, n_pulses, n_clusters,
      n_members;
} crv_timing;

/* The whole chain, one launch per stage: steps -> ADC records, pulses and clusters (plain FP64, the per-stage
 * functions in sequence on device-resident tables). The 128-sample records stay in the library's workspace
 * (crv_chain_records gives their device address); pulse and cluster tables are copied to the caller's buffers. */
int crv_chain_run(crv_ctx *ctx, const crv_steps *steps, crv_pulses *pulses, crv_clusters *clusters,
                  crv_timing *timing, crv_stream stream);
/* Copy the pulse and cluster tables of the last crv_chain_run again (after CRV_E_CAPACITY, into larger buffers). */
int crv_chain_fetch(crv_ctx *ctx, crv_pulses *pulses, crv_clusters *clusters, crv_stream stream);
/* Device addresses of the last chain run's waveform records: adc [n * window_samples] u16, label [n] i64. */
int crv_chain_records(crv_ctx *ctx, const uint16_t **adc, const int64_t **label, int64_t *n);

/* Bytes of device workspace the context holds now. Workspace slots grow to the largest call and are kept for the
 * next one; no reference counterpart (the reference's JAX twins leave device memory to XLA's allocator). */
int crv_ctx_workspace(crv_ctx *ctx, uint64_t *bytes);
/* Return all workspace to the device (after the device has drained), e.g. before the caller allocates large
 * buffers of its own. Forgets the last chain run's tables: crv_chain_fetch / crv_chain_records and the record
 * pointers of crv_batch_opts are invalid until the next chain call. No reference counterpart (see above). */
int crv_ctx_trim(crv_ctx *ctx, uint64_t *freed);

/* Options and diagnostics of the FP32 batched chain. */
typedef struct crv_batch_opts {
  int32_t grid;        /* persistent workgroups of 64 lanes of the SiPM kernels (0: all that fit on the device) */
  int32_t tile;        /* SiPMs per work-queue grab, a multiple of 64 (0: 64) */
  int32_t pulse_slots; /* initial pulse slots per reco waveform (0: 4); grown and re-run when one has more */
  int32_t sort_tasks;  /* SiPM order of the work queues: 0 as given (event, counter, SiPM); longest first, so that the
                        * lanes of a wave carry similar work, 1 by photon count, 2 by its octave and 3 by its quarter
                        * octave (SiPMs of one bucket in the given order, which keeps neighbours that share steps) */
  int64_t log_budget;  /* bytes of fired-pixel log per pass (0: half the device memory free at the call); a call whose
                        * log is larger runs its SiPMs in several passes of the photon and waveform kernels */
  /* written by the call: the window_samples-long ADC record of every cluster with a digi, in no particular order
   * (label -1: unused slot), in the library's workspace until the next call on the context */
  const uint16_t *record_adc;
  const int64_t *record_label;
  const int32_t *record_tick;
  int64_t n_records, n_waveforms;
  int32_t overlong, capped, pulse_slots_used, log_passes;
  int32_t pipeline; /* >= 2: the SiPMs run in at least this many passes, the waveform kernel of a pass beside the
                     * photon kernel of the next on a side stream (two fired-pixel log buffers); 0: one pass */
  int32_t _pad;
} crv_batch_opts;

/* The whole chain as the FP32 batched pipeline: steps -> pulses in three persistent work-queue kernels (photons and
 * SiPM pixels into a per-SiPM fired-pixel log; waveform records and pulse runs; pulse fits), no charge or digi
 * tables, then the coincidence finder. FP32 arithmetic; FP64
 * for absolute times and the 3 x 3 normal-equation solve of the fit. Same outputs and conventions as crv_chain_run
 * (pulse `wf` is the SiPM index of the call), equal to the FP64 chain statistically (same Philox streams, FP32
 * rounding). The reference runs this chain as chain.charges -> digis -> reco_waveforms -> clusters
 * (chain.py:151-362); this entry point has no single counterpart there. */
int crv_chain_run_batched(crv_ctx *ctx, const crv_steps *steps, crv_batch_opts *opts, crv_pulses *pulses,
                          crv_clusters *clusters, crv_timing *timing, crv_stream stream);

/* The FP32 batched chain captured into one HIP graph (a persistent, fixed-capacity launch sequence without host
 * synchronisation): the step grouping, the photon / SiPM / waveform work-queue kernels, the fits, the pulse table and
 * the coincidence finder of crv_chain_run_batched, for calls of at most max_steps steps. The reference has no
 * counterpart (its JAX twins dispatch the same chain asynchronously, reco/chain_jax.py:127-210, with fixed shapes,
 * response/device.py:110-180).
 *
 * crv_graph_create runs the chain once uncaptured on `sample` (a call typical of the workload) for the table sizes
 * per step, sizes every buffer for max_steps steps with (1 + margin) headroom, runs the fixed-capacity sequence
 * eagerly on the sample (it must equal the uncaptured chain's table sizes), then captures it. opts must give the
 * work-queue tile and sort order (e.g. from the autotuner); log passes and pipelining do not apply. The graph holds
 * its own device buffers until crv_graph_destroy. */
typedef struct crv_graph crv_graph;
int crv_graph_create(crv_ctx *ctx, const crv_steps *sample, int64_t max_steps, const crv_batch_opts *opts,
                     double margin, crv_graph **out);
/* Replay for `steps` (copied into the graph's input buffers on `stream`, then one hipGraphLaunch); outputs as
 * crv_chain_run_batched's (same tables, CRV_E_CAPACITY convention; crv_graph_fetch copies them again). A call that
 * exceeds a capacity (more steps than max_steps, or a table larger than planned) runs uncaptured through
 * crv_chain_run_batched instead, with the same results; crv_graph_stats counts those. The context's parameters must
 * be those of the capture (CRV_E_STATE otherwise); conditions, template and geometry uploads after the capture are
 * not seen by it. */
int crv_graph_launch(crv_graph *g, const crv_steps *steps, crv_pulses *pulses, crv_clusters *clusters,
                     crv_timing *timing, crv_stream stream);
int crv_graph_fetch(crv_graph *g, crv_pulses *pulses, crv_clusters *clusters, crv_stream stream);
/* launches, uncaptured fallbacks, and capacities[7]: steps, log chunks, records, runs, pulses, pulse slots per
 * waveform, whether the last launch fell back */
int crv_graph_stats(crv_graph *g, int64_t *launches, int64_t *fallbacks, int64_t *capacities);
int crv_graph_destroy(crv_graph *g);

#ifdef __cplusplus
}
#endif
#endif
