#!/usr/bin/env python3
"""CRV response chain benchmark on one MI355X: events/s from CrvSteps to coincidence clusters.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--events N] [--fp64-events M] [--workload cosmic|uneven|overlay]

A step is one pass of the chain over the whole synthetic event batch, inputs resident in HBM before timing, outputs
(pulse and cluster tables) copied into caller buffers inside the timed region. The headline is the FP32 batched chain
(crv_chain_run_batched) on N events; the FP64 plain chain (crv_chain_run, one launch per stage, bit-exact to the
oracle) is measured beside it on M events. The JSON line carries the dominant kernel's roofline (algorithmic bytes
per launch over its HIP-event duration on the chain's stream), the chain's byte contract per event, and the oracle
timed on this host's cores over event shards (a reported baseline).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mu2e-fastsim_amd"))

import numpy as np  # noqa: E402

METRIC = "CRV events/s (steps->coincidence clusters), FP32 batched vs FP64 plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
PROGRESS = False

# algorithmic bytes per table row (inputs read once, outputs written once)
STEP_B, RECORD_B, PULSE_B, CLUSTER_B, MEMBER_B = 32, 268, 32, 56, 4
# FP64 plain chain kernels
CHARGE_OUT_B = 22      # time f64, voltage f64, step row i32, pixel i16 written by the SiPM kernel
CHARGE_IN_B = 16       # time and charge read by the digitisation kernel
DIGI_SLOT_B = 16       # record, sample, start of a digi
SAMPLE_B = 2           # one ADC word read by the pulse kernel
WAVEFORM_META_B = 40   # start, length, offset, calibration of a reco waveform
PULSE_IN_B = 40        # event, bar, sipm, time, pe, seq read by the finder
# FP32 batched chain kernels
SIPM_STEP_B = 32       # per step read by kb_sipm: z, t f64 and the photon counts of its 4 SiPMs i32
FIRE_B = 8             # fired-pixel log entry written by kb_sipm: time offset f32, overvoltage f32
FIRE_IN_B = 8          # log entry read by the waveform kernel (kb_wrec): time offset, overvoltage
RUN_B = 32             # reco-waveform descriptor: record, first sample, samples, run number i32; SiPM, start TDC i64
SLOT_B = 48            # pulse slot written by kb_fit: time, location f64; 7 x f32; flags u32


def generate(workload, n, seed):
    from crvhip import synth as SY
    if workload == "cosmic":
        return SY.cosmic_events(n, seed=seed)
    if workload == "uneven":
        return SY.uneven_events(n, seed=seed)
    if workload == "overlay":
        return SY.overlay_events(n, seed=seed)
    raise SystemExit(f"unknown workload {workload}")


def _cpu_shard(args):
    steps, seed = args
    from crvhip import config as CF, detector as DT, template as TP
    from oracle import chain as OC
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    t0 = time.perf_counter()
    out = OC.run(steps, cond, DT.sector_geom(), cfg, seed, TP.load(cfg.waveform))
    return len(out["clusters"]["cl_event"]), time.perf_counter() - t0


def cpu_baseline(S, n_events, seed, per_worker, workers):
    """The oracle chain on `workers` processes over contiguous event shards of the same steps (fork, no GPU). The
    pool is started and every worker has run one event before the clock starts, so start-up is not timed."""
    import multiprocessing as mp
    ev = S["evkey"]
    shards = []
    for w in range(workers):
        a, b = w * per_worker, (w + 1) * per_worker
        if b > n_events:
            break
        lo, hi = np.searchsorted(ev, a), np.searchsorted(ev, b)
        shards.append(({k: v[lo:hi] for k, v in S.items()}, seed))
    one = np.searchsorted(ev, 1)
    warm = [({k: v[:one] for k, v in S.items()}, seed)] * len(shards)
    ctx = mp.get_context("fork")
    with ctx.Pool(len(shards)) as pool:
        pool.map(_cpu_shard, warm, chunksize=1)
        t0 = time.perf_counter()
        res = pool.map(_cpu_shard, shards, chunksize=1)
        dt = time.perf_counter() - t0
    n = len(shards) * per_worker
    return {"value": n / dt, "unit": "events/s", "cores": len(shards), "kind": "port",
            "sample": f"{n} events ({len(shards)} processes, one shard of {per_worker} contiguous events each, of the "
                      f"same synthetic batch), oracle chain steps->clusters (FP64), {dt:.1f} s wall after pool "
                      f"start-up; slowest shard {max(r[1] for r in res):.1f} s"}


class Outputs:
    """Caller-side pulse and cluster buffers (torch device tensors) of the given capacities."""

    def __init__(self, L, torch, cap_p, cap_c):
        cap_p, cap_c = max(int(cap_p), 1), max(int(cap_c), 1)
        b = {k: torch.empty(cap_p, dtype=torch.int64, device="cuda") for k in ("wf", "seq", "ev")}
        b["chan"] = torch.empty(cap_p, dtype=torch.int32, device="cuda")
        for k in L.PULSE_F64:
            b[k] = torch.empty(cap_p, dtype=torch.float64, device="cuda")
        for k in L.PULSE_U8:
            b[k] = torch.empty(cap_p, dtype=torch.uint8, device="cuda")
        c = {"cl_event": torch.empty(cap_c, dtype=torch.int64, device="cuda"),
             "cl_mem_off": torch.empty(cap_c + 1, dtype=torch.int64, device="cuda"),
             "mem_idx": torch.empty(cap_p, dtype=torch.int64, device="cuda")}
        for k in ("cl_type", "cl_poscalc"):
            c[k] = torch.empty(cap_c, dtype=torch.int32, device="cuda")
        c["cl_sidehits"] = torch.empty(2 * cap_c, dtype=torch.int32, device="cuda")
        for k in L.CLUSTER_F64:
            m = 3 if k == "cl_pos" else 2 if k in ("cl_sidepe", "cl_sidetime") else 1
            c[k] = torch.empty(m * cap_c, dtype=torch.float64, device="cuda")
        self.keep = (b, c)
        self.P = L.Pulses(*[L.ptr(b[k]) for k in ("wf", "seq", "ev", "chan") + L.PULSE_F64 + L.PULSE_U8], cap_p, 0)
        self.C = L.Clusters(L.ptr(c["cl_event"]), L.ptr(c["cl_type"]), L.ptr(c["cl_poscalc"]),
                            L.ptr(c["cl_sidehits"]), *[L.ptr(c[k]) for k in L.CLUSTER_F64], L.ptr(c["cl_mem_off"]),
                            cap_c, 0, L.ptr(c["mem_idx"]), cap_p, 0)


def chunks_of(evkey, n_ev, chunk):
    """(first step, steps) of each chunk of `chunk` events (the whole batch when chunk is 0); evkey is the steps'
    event key, a numpy array or a device tensor."""
    if not chunk or chunk >= n_ev:
        return [(0, len(evkey))]
    edges = np.arange(0, n_ev + chunk, chunk).clip(max=n_ev)
    if isinstance(evkey, np.ndarray):
        bounds = np.searchsorted(evkey, edges)
    else:
        import torch
        bounds = torch.searchsorted(evkey, torch.as_tensor(edges, dtype=evkey.dtype, device=evkey.device)).cpu().numpy()
    return [(int(a), int(b - a)) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]


AUTOTUNE_EVENTS = 200_000  # events of the sample the work-queue autotuner runs its candidates on
GRAPH_MARGIN = 0.3         # headroom of a captured graph's tables over the first chunk's, per step


def measure(precision, ds, evkey, n_ev, seed, steps, warmup, batch, chunk=0, autotune=False, tables=False,
            graph=False):
    """Time `steps` chain passes over the batch (chunk by chunk, steps resident in HBM); returns (events/s, ms/step,
    mean timing fields summed over chunks, last counts summed over chunks, extra). With `tables` (one chunk) extra
    holds the last pass's pulse columns ev, chan, time, pe on the host. With `graph` (FP32) every chunk replays one
    captured graph of the chain (crv_graph_launch), sized on the first chunk for the largest chunk's steps."""
    import torch
    from crvhip import _lib as L, config as CF, detector as DT, template as TP
    cfg = CF.ChainConfig(precision=precision)
    ctx = L.Context(cfg, DT.nominal_conditions(), DT.sector_geom(), TP.load(cfg.waveform), 0)
    ctx.set_params(cfg, seed, ds.evkey_hi, 32 if precision == "fp32" else 64)
    parts = chunks_of(evkey, n_ev, chunk)
    base = ds.struct()
    sizes = (4, 4, 8, 8, 8)
    structs = [L.Steps(*[getattr(base, f) + a * sz for f, sz in zip(("event", "counter", "edep", "z", "t"), sizes)],
                       n) for a, n in parts]
    tm = L.Timing()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    tuned = None
    if autotune and precision == "fp32":  # tile / sort order of the work queues, measured on the batch's first events
        from crvhip import autotune as AT
        n_tune = min(n_ev, AUTOTUNE_EVENTS, chunk or n_ev)  # never more events than one timed chunk holds
        a, n = chunks_of(evkey, n_ev, n_tune)[0]
        sample = L.Steps(*[getattr(base, f) + a * sz for f, sz in zip(("event", "counter", "edep", "z", "t"), sizes)],
                         n)
        chosen, rep = AT.tune(ctx, sample, n_tune, torch.cuda.get_device_name(0), stream=stream, cache=False)
        batch = {**batch, **chosen}
        tuned = {"sample_events": n_tune, "chosen": chosen,
                 "median_ms": {f"tile {c['opts']['tile']} sort {c['opts']['sort_tasks']}":
                               round(sorted(c["ms"])[len(c["ms"]) // 2], 2) for c in rep["candidates"]}}
        if PROGRESS:
            print(f"[bench] autotune: {chosen}", file=sys.stderr, flush=True)
    opts = L.BatchOpts(**batch)
    G = None
    if graph and precision == "fp32":
        if not opts.tile:
            opts.tile = 64
        t0 = time.perf_counter()
        G = L.Graph(ctx, structs[0], max(st.n for st in structs), opts, GRAPH_MARGIN)
        graph_s = time.perf_counter() - t0

    def call(st, out):
        if G is not None:
            return G.launch(st, out.P, out.C, tm, stream)
        if precision == "fp32":
            return ctx.lib.crv_chain_run_batched(ctx.h, ctypes.byref(st), ctypes.byref(opts), ctypes.byref(out.P),
                                                 ctypes.byref(out.C), ctypes.byref(tm), stream)
        return ctx.lib.crv_chain_run(ctx.h, ctypes.byref(st), ctypes.byref(out.P), ctypes.byref(out.C),
                                     ctypes.byref(tm), stream)

    # sizing pass (also the first warmup): zero-capacity outputs report the sizes they need
    need_p = need_c = need_m = 1
    empty = Outputs(L, torch, 1, 1)
    empty.P.capacity = empty.C.capacity = empty.C.mem_capacity = 0
    for st in structs:
        rc = call(st, empty)
        if rc not in (0, L.CRV_E_CAPACITY):
            ctx.check(rc)
        need_p, need_c, need_m = max(need_p, empty.P.n), max(need_c, empty.C.n), max(need_m, empty.C.n_mem)
    try:
        out = Outputs(L, torch, need_p, max(need_c, need_m))
    except torch.OutOfMemoryError:
        # the library's cached workspace of the sizing pass holds the device: give it back (the warmup pass below
        # allocates it again, outside the timed region)
        ctx.trim()
        torch.cuda.empty_cache()
        out = Outputs(L, torch, need_p, max(need_c, need_m))

    def step():
        acc = {}
        for i, st in enumerate(structs):
            ctx.check(call(st, out))
            if PROGRESS:
                print(f"[bench] {precision} chunk {i + 1}/{len(structs)}: {tm.total_ms:.1f} ms", file=sys.stderr,
                      flush=True)
            for f, _ in L.Timing._fields_:
                acc[f] = acc.get(f, 0) + getattr(tm, f)
        return acc

    for _ in range(warmup):  # full passes into the caller's buffers after the sizing pass
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = [step() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    avg = {k: float(np.mean([r[k] for r in runs])) for k in runs[0]}
    extra = {"n_waveforms_batched": opts.n_waveforms, "pulse_slots_used": opts.pulse_slots_used} \
        if precision == "fp32" else {}
    extra["chunks"] = len(structs)
    if tuned:
        extra["autotune"] = tuned
    if G is not None:
        extra["graph"] = {**G.stats(), "create_s": round(graph_s, 2), "margin": GRAPH_MARGIN}
        G.close()
    if tables:
        assert len(structs) == 1, "pulse tables of a one-chunk run only"
        extra["pulses"] = {k: out.keep[0][k][:out.P.n].cpu().numpy() for k in ("ev", "chan", "time", "pe")}
    ctx.close()
    del out, empty
    torch.cuda.empty_cache()
    return n_ev * steps / dt, dt / steps * 1e3, avg, runs[-1], extra


def pulse_errors(P32, P64):
    """FP32 pulse time and PE against the FP64 chain's on the same events: the pulses of each (event, channel) whose
    pulse count is the same in both tables, paired in order (both tables are in (event, channel, time) order)."""
    k32 = P32["ev"].astype(np.int64) * 32768 + P32["chan"]
    k64 = P64["ev"].astype(np.int64) * 32768 + P64["chan"]
    u32, n32 = np.unique(k32, return_counts=True)
    u64, n64 = np.unique(k64, return_counts=True)
    both, i32, i64 = np.intersect1d(u32, u64, assume_unique=True, return_indices=True)
    keys = both[n32[i32] == n64[i64]]
    m32, m64 = np.isin(k32, keys), np.isin(k64, keys)
    dt = np.abs(P32["time"][m32] - P64["time"][m64])
    dpe = np.abs(P32["pe"][m32] - P64["pe"][m64]) / np.maximum(np.abs(P64["pe"][m64]), 1e-9)
    q = lambda a, f: float(np.quantile(a, f)) if len(a) else None  # noqa: E731
    return {"note": "pulses of the (event, channel) pairs with the same pulse count in both chains, paired in time "
                    "order; |dt| in ns, |dPE| relative",
            "paired_fraction": float(m64.sum() / max(len(k64), 1)),
            "pe_sum_rel": float((P32["pe"].sum() - P64["pe"].sum()) / max(abs(P64["pe"].sum()), 1e-9)),
            "dt_median_ns": q(dt, 0.5), "dt_p99_ns": q(dt, 0.99), "dt_max_ns": float(dt.max()) if len(dt) else None,
            "dpe_median": q(dpe, 0.5), "dpe_p99": q(dpe, 0.99)}


def kernels_fp64(avg, cnt):
    S_, W_, P_, C_, M_ = (cnt["n_steps"], cnt["n_records"], cnt["n_pulses"], cnt["n_clusters"], cnt["n_members"])
    return {
        "k_sipm": (avg["k_sipm_ms"], STEP_B * S_ + CHARGE_OUT_B * cnt["n_charges"]),
        "k_wf_synth": (avg["k_digi_ms"], CHARGE_IN_B * cnt["n_charges"] + RECORD_B * W_ + DIGI_SLOT_B * cnt["n_digis"]),
        "k_pulses": (avg["k_pulse_ms"], SAMPLE_B * 12 * cnt["n_digis"] + WAVEFORM_META_B * cnt["n_waveforms"]
                     + PULSE_B * P_),
        "k_coinc": (avg["k_coinc_ms"], PULSE_IN_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_),
    }


def kernels_fp32(avg, cnt, fires):
    S_, W_, R_, P_, C_, M_ = (cnt["n_steps"], cnt["n_records"], cnt["n_waveforms"], cnt["n_pulses"],
                              cnt["n_clusters"], cnt["n_members"])
    return {
        "kb_sipm": (avg["k_sipm_ms"], SIPM_STEP_B * S_ + FIRE_B * fires),
        "kb_wrec": (avg["k_digi_ms"], FIRE_IN_B * fires + RECORD_B * W_ + RUN_B * R_),
        "kb_fit": (avg["k_pulse_ms"], RUN_B * R_ + SAMPLE_B * 24 * R_ + SLOT_B * P_),
        "k_coinc": (avg["k_coinc_ms"], PULSE_IN_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_),
    }


# HBM bytes per event actually moved by a kernel on the cosmic workload: rocprofv3 FETCH_SIZE + WRITE_SIZE passes
# over 1M events, each scaled by the factor that makes a plain 1 GiB device copy read and write 1 GiB
# (profiles/chain_fp32_cosmic1m_traffic.txt); used for roofline.traffic of the same workload only
PMC_BYTES_PER_EVENT = {"cosmic": {"kb_sipm": 90919.7, "kb_wrec": 35395.0, "kb_fit": 24742.5}}
# VALU issue and active lanes of the chain's kernels on the cosmic workload, from rocprofv3 SQ counters over 1M events
# (profiles/chain_fp32_cosmic1m_valu.txt): busy = 4 SQ_ACTIVE_INST_VALU / 1024 SIMDs / GRBM_GUI_ACTIVE cycles per XCD,
# lanes = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)
PMC_VALU = {"cosmic": {"kb_sipm": {"busy": 0.795, "lanes": 0.346, "valu_per_event": 35907},
                       "kb_wrec": {"busy": 0.792, "lanes": 0.427, "valu_per_event": 10794},
                       "kb_fit": {"busy": 1.088, "lanes": 0.664, "valu_per_event": 10209},
                       "k_coinc_wave<64>": {"busy": 0.569, "lanes": 0.539, "valu_per_event": 7175}}}


def roofline(kernels, workload=None, events=0, per_call=1):
    """The dominant kernel's algorithmic bytes over its HIP-event time (kernel times and bytes are summed over the
    `per_call` calls of a pass; per-launch figures divide by them; `events` per call)."""
    dom = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes = kernels[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    per_event = PMC_BYTES_PER_EVENT.get(workload, {}).get(dom)
    out = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": per_event * events if per_event else None,
           "bytes_per_launch": kbytes / per_call, "ms_per_launch": kms / per_call}
    valu = PMC_VALU.get(workload, {}).get(dom)
    if valu:  # the kernel is issue-bound: its VALU issue and active-lane fraction from the counter profile
        out["valu"] = {**valu, "source": "rocprofv3 SQ counters, same workload (profiles/chain_fp32_cosmic1m_valu.txt)"}
    return out


LK_STEP_B = 128        # per step read by the lookup optics: event, bar i32; t0, t1, edep, len, p0, p1, mass, charge,
#                        start and end positions f64
PHOTON_B = 32          # per photon written by the lookup optics (label, time, step row f64/i64; step ordinal, rank i32)
SIPM_CHARGE_B = 40     # per charge written by the SiPM stage (label, time, charge, pe, photon)


def lookup_leg(n_events, seed, steps):
    """The stock lookup-table optics on the synthetic detector (one made-up v6 table, crvhip.synth_lookup): the cosmic
    workload's steps as CrvStep rows through crv_lookup_photons_run (photons at the SiPMs, FP64), then crv_sipm_run
    (charges), inputs resident in HBM; each call is timed on the host (the ABI returns once its stream has finished).
    A sizing pass and a warmup precede the `steps` timed passes."""
    import torch
    from crvhip import _lib as L, chain as CH, config as CF, detector as DT, photons as PH, synth as SY
    from crvhip import synth_lookup as SL, template as TP
    S = SY.cosmic_events(n_events, seed=seed)
    cols, pdg = SL.crv_steps(S)
    del S
    T = SL.table()
    G, _ = DT.synthetic_crv_geometry()
    cfg = CF.ChainConfig(response=SL.response_config())
    ctx = L.Context(cfg, DT.nominal_conditions(), None, TP.load(cfg.waveform), 0)
    PH.upload(ctx, {T.name: T}, G, cfg.response)
    ls = PH.LookupSteps(CH.event_key(cols), cols, pdg)
    ctx.set_params(cfg, seed + 1, ls.evkey_hi)
    pb, po = PH.device_photons(ctx, ls)  # sizing pass: the photon buffers at the size the call needs
    nph = po.n
    pl = L.PhotonList(L.ptr(pb["label"]), L.ptr(pb["time"]), L.ptr(pb["photon"]), L.ptr(pb["j"]), L.ptr(pb["rank"]),
                      nph)
    sc = cfg.sipm
    nz = L.SipmNoise(0.0, 0.0, 0.0, float(sc.trap0_lifetime), float(sc.trap1_lifetime), 0.0, float(cfg.window[1]),
                     None, 0)
    cap = 2 * nph + 4096
    for _ in range(2):
        qb = {k: torch.empty(cap, dtype=dt, device="cuda") for k, dt in (
            ("label", torch.int64), ("time", torch.float64), ("charge", torch.float64), ("pe", torch.float64),
            ("photon", torch.int64))}
        qo = L.Charges(*[L.ptr(qb[k]) for k in ("label", "time", "charge", "pe", "photon")], cap, 0, 0, 0)
        rc = ctx.lib.crv_sipm_run(ctx.h, ctypes.byref(pl), ctypes.byref(nz), ctypes.byref(qo), None)
        if rc == L.CRV_E_CAPACITY:
            cap = qo.n
            continue
        ctx.check(rc)
        break
    st = ls.struct()
    t_ph, t_si = [], []
    for it in range(steps + 1):  # a warmup, then the timed passes
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.check(ctx.lib.crv_lookup_photons_run(ctx.h, ctypes.byref(st), ctypes.byref(po), None))
        t1 = time.perf_counter()
        ctx.check(ctx.lib.crv_sipm_run(ctx.h, ctypes.byref(pl), ctypes.byref(nz), ctypes.byref(qo), None))
        t2 = time.perf_counter()
        if it:
            t_ph.append(t1 - t0)
            t_si.append(t2 - t1)
    assert po.n == nph, "the photons of a pass changed"
    ph_s, si_s = float(np.mean(t_ph)), float(np.mean(t_si))
    ph_bytes = LK_STEP_B * ls.n + PHOTON_B * nph
    si_bytes = PHOTON_B * nph + SIPM_CHARGE_B * qo.n
    out = {"value": n_events / (ph_s + si_s), "unit": "events/s", "events": n_events, "steps": steps,
           "ms_per_step": (ph_s + si_s) * 1e3, "dtype": "f64",
           "config": {"workload": "cosmic events as CrvStep rows, one synthetic v6 lookup table "
                                  f"({sum(t.delays.nbytes for t in T.tables.values()) >> 20} MiB of delay weights)",
                      "stages": "crv_lookup_photons_run -> crv_sipm_run"},
           "stage_ms": {"photons": ph_s * 1e3, "sipm": si_s * 1e3},
           "roofline": {"bound": "hbm", "kernel": "lookup photons stage (count, scan, fill)",
                        "achieved": ph_bytes / ph_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ph_bytes / ph_s / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "bytes_per_launch": ph_bytes, "ms_per_launch": ph_s * 1e3},
           "sipm_stage": {"achieved": si_bytes / si_s / 1e9, "unit": "GB/s", "bytes_per_launch": si_bytes},
           "per_event": {"steps": ls.n / n_events, "photons": nph / n_events, "charges": qo.n / n_events}}
    ctx.close()
    del pb, qb, ls
    torch.cuda.empty_cache()
    return out


def workload_leg(workload, n_events, chunk, seed, steps, warmup, batch, autotune, graph=False):
    """One extra workload (device-generated, resident in HBM) through the FP32 batched chain in calls of `chunk`
    events: events/s over `steps` timed passes after the sizing pass, with its own dominant-kernel roofline."""
    import torch
    from crvhip import synth_gpu as SG
    t0 = time.perf_counter()
    Sd = SG.events(workload, n_events, seed)
    ds = SG.device_steps(Sd)
    del Sd
    torch.cuda.synchronize()
    torch.cuda.empty_cache()  # the generator's temporaries: cached by torch, they would starve the library's hipMalloc
    gen_s = time.perf_counter() - t0
    v, ms, avg, cnt, ex = measure("fp32", ds, ds.event, n_events, seed + 1, steps, warmup, batch, chunk, autotune,
                                  graph=graph)
    tuned = ex.pop("autotune", None)
    k = kernels_fp32(avg, cnt, cnt["n_charges"])
    S_, W_, P_, C_, M_ = (cnt["n_steps"], cnt["n_records"], cnt["n_pulses"], cnt["n_clusters"], cnt["n_members"])
    ncalls = ex["chunks"]
    b_event = (STEP_B * S_ + RECORD_B * W_ + PULSE_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_) / max(n_events, 1)
    out = {"value": v, "unit": "events/s", "events": n_events, "chunk": chunk or n_events, "calls": ncalls,
           "steps": steps, "ms_per_step": ms, "dtype": "f32", "generator": "device (torch)",
           "generate_s": round(gen_s, 2), "options": {**batch, **(tuned["chosen"] if tuned else {})},
           "autotune": tuned, "graph": ex.get("graph"),
           "roofline": roofline(k, workload, chunk or n_events, per_call=ncalls),
           "chain_roofline": {"bytes_per_event": b_event, "achieved": b_event * v / 1e9, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": b_event * v / 1e9 / HBM_PEAK_GBS},
           "kernel_ms": {kk: vv[0] for kk, vv in k.items()},
           "stage_ms": {kk: avg[kk] for kk in ("charges_ms", "digis_ms", "pulses_ms", "clusters_ms", "total_ms")},
           "per_event": {kk[2:]: cnt[kk] / n_events for kk in cnt if kk.startswith("n_")}}
    del ds
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=1_000_000, help="events of the FP32 batched chain (headline)")
    ap.add_argument("--fp64-events", type=int, default=1_000_000, help="events of the FP64 plain chain (0: skip)")
    ap.add_argument("--fp64-steps", type=int, default=3, help="timed passes of the FP64 plain chain")
    ap.add_argument("--fp64-chunk", type=int, default=250_000,
                    help="events per FP64 chain call (its charge and digi tables of 1M events exceed the device)")
    ap.add_argument("--error-events", type=int, default=100_000,
                    help="events on which the FP32 pulses are paired with the FP64 pulses one by one")
    ap.add_argument("--uneven-events", type=int, default=10_000_000, help="uneven-mixture leg (0: skip)")
    ap.add_argument("--uneven-chunk", type=int, default=500_000)
    ap.add_argument("--no-graph", action="store_true",
                    help="uneven leg: run every chunk through crv_chain_run_batched instead of one captured graph")
    ap.add_argument("--overlay-events", type=int, default=1_000_000, help="high-occupancy overlay leg (0: skip)")
    ap.add_argument("--overlay-chunk", type=int, default=10_000,
                    help="overlay events per call (the 1M overlay events' steps alone take 89 GB of HBM)")
    ap.add_argument("--lookup-events", type=int, default=100_000,
                    help="lookup-table optics leg: cosmic events as CrvStep rows through the photon and SiPM stages "
                         "(0: skip)")
    ap.add_argument("--workload", default="cosmic")
    ap.add_argument("--graph", action="store_true", help="headline: replay one captured graph per chunk")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--pipeline", type=int, default=0,
                    help="FP32 chain: at least this many SiPM passes, waveforms of one beside photons of the next")
    ap.add_argument("--sort", type=int, default=1, choices=(0, 1, 2, 3),
                    help="work-queue SiPM order: 0 as given, longest first by photon count (1), its octave (2) or "
                         "quarter octave (3)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --tile / --sort as given (with --tile 0 the work-queue tile and order are autotuned)")
    ap.add_argument("--chunk", type=int, default=0, help="events per chain call (0: the whole batch in one call)")
    ap.add_argument("--gen", default="auto", choices=("auto", "host", "device"),
                    help="synthetic input generator: numpy on the host or torch on the device (auto: device for "
                         "uneven / overlay and for more than 2M events)")
    ap.add_argument("--cpu-events", type=int, default=12_000, help="events of the oracle CPU baseline (0: skip)")
    ap.add_argument("--cpu-workers", type=int, default=0, help="oracle processes (0: every core this process may use)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer three times (the reference kernel for HBM counters)")
    ap.add_argument("--progress", action="store_true", help="a line on stderr per chain call")
    args = ap.parse_args()
    global PROGRESS
    PROGRESS = args.progress
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (events would shard by contiguous "
              "event ranges, one process per GPU). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # inputs: numpy on the host (the oracle's own generator) for the cosmic batch, torch on the device for the large
    # uneven / overlay batches; the CPU baseline runs first, before anything touches the GPU
    gen_device = args.gen == "device" or (args.gen == "auto" and (args.workload != "cosmic" or args.events > 2_000_000))
    S = None if gen_device else generate(args.workload, args.events, args.seed)
    cpu = None
    if not args.no_cpu and args.cpu_events:
        # one process per core of this GPU's share of the host (16 cores per GPU on the measurement boxes)
        workers = args.cpu_workers or min(len(os.sched_getaffinity(0)), 16)
        per_worker = max(1, args.cpu_events // workers)
        n_cpu = per_worker * workers
        sample = S if S is not None and n_cpu <= args.events else generate(args.workload, n_cpu, args.seed)
        cpu = cpu_baseline(sample, n_cpu, args.seed + 1, per_worker, workers)
        if sample is not S:
            cpu["sample"] += " (host-generated events of the same workload; the GPU batch is generated on the device)"
        progress(f"cpu baseline: {cpu['value']:.0f} events/s on {cpu['cores']} processes")

    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    from crvhip.chain import DeviceSteps
    if gen_device:
        from crvhip import synth_gpu as SG
        Sd = SG.events(args.workload, args.events, args.seed)
        ds = SG.device_steps(Sd)
        del Sd
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        evkey = ds.event
    else:
        ds = DeviceSteps(S["evkey"], S["counter"], S["edep"], S["z"], S["t"])
        evkey = S["evkey"]
    batch = {"grid": args.grid, "tile": args.tile, "sort_tasks": args.sort, "pipeline": args.pipeline}
    autotune = args.tile == 0 and not args.no_autotune
    v32, ms32, avg32, cnt32, ex32 = measure("fp32", ds, evkey, args.events, args.seed + 1, args.steps, args.warmup,
                                            batch, args.chunk, autotune, graph=args.graph)
    progress(f"fp32 {args.workload} {args.events}: {v32:.0f} events/s")
    tuned = ex32.pop("autotune", None)
    if tuned:
        batch = {**batch, **tuned["chosen"]}
    k32 = kernels_fp32(avg32, cnt32, cnt32["n_charges"])
    fp64 = None
    if args.fp64_events and args.workload == "cosmic":
        n64 = min(args.fp64_events, args.events)
        if n64 == args.events:
            ds64, ev64 = ds, evkey
        else:
            (_, cut), = chunks_of(evkey, args.events, n64)[:1]
            ev64 = evkey[:cut] if isinstance(evkey, np.ndarray) else evkey[:cut].cpu().numpy().astype(np.int64)
            ds64 = DeviceSteps(ev64, ds.counter[:cut].cpu().numpy(), ds.edep[:cut].cpu().numpy(),
                               ds.z[:cut].cpu().numpy(), ds.t[:cut].cpu().numpy())
        v64, ms64, avg64, cnt64, ex64 = measure("fp64", ds64, ev64, n64, args.seed + 1, args.fp64_steps, args.warmup,
                                                batch, args.fp64_chunk)
        progress(f"fp64 {n64}: {v64:.0f} events/s")
        del ds64
        # FP32 counts against the FP64 chain on the same events (same seed, so the same Philox streams): from the
        # headline run when it covered the same events, else from an FP32 pass over them
        c32s = cnt32 if n64 == args.events else measure("fp32", *_sub(ds, evkey, args.events, n64), n64,
                                                        args.seed + 1, 1, 0, batch)[3]
        fp32_err = {k: (c32s[f"n_{k}"] - cnt64[f"n_{k}"]) / max(cnt64[f"n_{k}"], 1)
                    for k in ("photons", "charges", "pulses", "clusters", "members")}
        # pulse by pulse on the first error-events events
        ne = min(args.error_events, n64)
        dse, eve = _sub(ds, evkey, args.events, ne)
        _, _, _, _, exe64 = measure("fp64", dse, eve, ne, args.seed + 1, 1, 0, batch, tables=True)
        _, _, _, _, exe32 = measure("fp32", dse, eve, ne, args.seed + 1, 1, 0, batch, tables=True)
        del dse
        fp32_pulse_err = {"events": ne, **pulse_errors(exe32.pop("pulses"), exe64.pop("pulses"))}
        k64 = kernels_fp64(avg64, cnt64)
        fp64 = {"value": v64, "unit": "events/s", "events": n64, "chunk": min(args.fp64_chunk or n64, n64),
                "calls": ex64["chunks"], "steps": args.fp64_steps, "ms_per_step": ms64,
                "dtype": "f64", "speedup_fp32_over_fp64": v32 / v64,
                "roofline": roofline(k64, per_call=ex64["chunks"]),
                "fp32_rel_count_error": {"events": n64, "note": "photons, fired pixels (integer PE), pulses, "
                                         "clusters, cluster members: (FP32 - FP64) / FP64 on the same events", **fp32_err},
                "fp32_pulse_error": fp32_pulse_err,
                "kernel_ms": {k: v[0] for k, v in k64.items()},
                "stage_ms": {k: avg64[k] for k in ("charges_ms", "digis_ms", "waveforms_ms", "pulses_ms",
                                                   "clusters_ms", "total_ms")},
                "per_event": {k[2:]: cnt64[k] / n64 for k in cnt64 if k.startswith("n_")}}
    del ds
    torch.cuda.empty_cache()
    legs = {}
    if args.workload == "cosmic":
        if args.uneven_events:
            legs["uneven"] = workload_leg("uneven", args.uneven_events, args.uneven_chunk, args.seed, 1, 0,
                                          {**batch, "tile": args.tile, "sort_tasks": args.sort}, autotune,
                                          graph=not args.no_graph)
            progress(f"uneven {args.uneven_events}: {legs['uneven']['value']:.0f} events/s")
        if args.overlay_events:
            legs["overlay"] = workload_leg("overlay", args.overlay_events, args.overlay_chunk, args.seed, 1, 0,
                                           {**batch, "tile": args.tile, "sort_tasks": args.sort}, autotune)
            progress(f"overlay {args.overlay_events}: {legs['overlay']['value']:.0f} events/s")
        if args.lookup_events:
            legs["lookup"] = lookup_leg(args.lookup_events, args.seed, 3)
            progress(f"lookup {args.lookup_events}: {legs['lookup']['value']:.0f} events/s")
    S_, W_, P_, C_, M_ = (cnt32["n_steps"], cnt32["n_records"], cnt32["n_pulses"], cnt32["n_clusters"],
                          cnt32["n_members"])
    b_event = (STEP_B * S_ + RECORD_B * W_ + PULSE_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_) / args.events
    line = {
        "metric": METRIC, "value": v32, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload} events, FP32 batched chain (3 fused kernels + finder)",
                   "events": args.events, "chunk": args.chunk or args.events, "seed": args.seed,
                   "generator": "device (torch)" if gen_device else "host (numpy)",
                   "precision": "fp32 (fp64 times and 3x3 solve)",
                   "chain": "steps->photons->SiPM->waveforms->pulses->clusters", "grid": args.grid,
                   "tile": batch["tile"] or 64, "sort_tasks": batch["sort_tasks"],
                   "autotune": tuned},
        "roofline": roofline(k32, args.workload, args.chunk or args.events),
        "chain_roofline": {"bytes_per_event": b_event, "achieved": b_event * v32 / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": b_event * v32 / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "fp64_plain": fp64,
        "workloads": legs,
        "stage_ms": {k: avg32[k] for k in ("charges_ms", "digis_ms", "pulses_ms", "clusters_ms", "total_ms")},
        "kernel_ms": {k: v[0] for k, v in k32.items()},
        "per_event": {k[2:]: cnt32[k] / args.events for k in cnt32 if k.startswith("n_")},
        "batched": ex32,
    }
    print(json.dumps(line))


def _sub(ds, evkey, n_ev, n):
    """Device steps of the first n events of a batch (copies) and their event keys."""
    from crvhip.chain import DeviceSteps
    (_, cut), = chunks_of(evkey, n_ev, n)[:1]
    ev = evkey[:cut] if isinstance(evkey, np.ndarray) else evkey[:cut].cpu().numpy().astype(np.int64)
    return DeviceSteps(ev, ds.counter[:cut].cpu().numpy(), ds.edep[:cut].cpu().numpy(), ds.z[:cut].cpu().numpy(),
                       ds.t[:cut].cpu().numpy()), ev


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
