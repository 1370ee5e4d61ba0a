#!/usr/bin/env python3
"""CRV response chain benchmark on one MI355X: events/s from CrvSteps to coincidence clusters.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--events N] [--workload cosmic|uneven|overlay]

A step is one pass of the chain over the whole synthetic event batch (inputs resident in HBM before timing). The JSON
line carries the dominant kernel's roofline (algorithmic bytes per launch over its HIP-event duration), the chain's
byte-contract roofline (32 S + 268 W + 32 P + 56 C + 4 M bytes per event), and the oracle timed on this host's
cores over event shards (a reported baseline).
"""
import argparse
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

# algorithmic bytes per table row (inputs read once, outputs written once)
STEP_B, RECORD_B, PULSE_B, CLUSTER_B, MEMBER_B = 32, 268, 32, 56, 4
CHARGE_OUT_B = 22      # time f64, voltage f64, step row i32, pixel i16 written by the SiPM kernel
CHARGE_IN_B = 16       # time and charge read by the digitisation kernel
DIGI_SLOT_B = 16       # record, sample, start of a digi
SAMPLE_B = 2           # one ADC word read by the pulse kernel
WAVEFORM_META_B = 40   # start, length, offset, calibration of a reco waveform
PULSE_IN_B = 40        # event, bar, sipm, time, pe, seq read by the finder


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
    out = OC.run(steps, cond, DT.sector_geom(), cfg, seed, TP.load(cfg.waveform))
    return len(out["clusters"]["cl_event"])


def cpu_baseline(S, n_events, seed, per_worker, workers):
    """The oracle chain on `workers` processes over contiguous event shards of the same steps (fork, no GPU)."""
    import multiprocessing as mp
    ev = S["evkey"]
    shards = []
    for w in range(workers):
        a, b = w * per_worker, (w + 1) * per_worker
        if b > n_events:
            break
        lo, hi = np.searchsorted(ev, a), np.searchsorted(ev, b)
        shards.append(({k: v[lo:hi] for k, v in S.items()}, seed))
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(len(shards)) as pool:
        pool.map(_cpu_shard, shards)
    dt = time.perf_counter() - t0
    n = len(shards) * per_worker
    return {"value": n / dt, "unit": "events/s", "cores": len(shards), "kind": "port",
            "sample": f"{n} events ({len(shards)} shards of {per_worker} contiguous events of the same synthetic "
                      f"batch), oracle chain steps->clusters, {dt:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=100_000)
    ap.add_argument("--workload", default="cosmic")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-per-worker", type=int, default=40)
    ap.add_argument("--cpu-workers", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (events would shard by contiguous "
              "event ranges, one process per GPU). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from crvhip import _lib as L, config as CF, detector as DT, template as TP
    from crvhip.chain import DeviceSteps
    S = generate(args.workload, args.events, args.seed)

    cpu = None
    if not args.no_cpu:
        workers = args.cpu_workers or min(len(os.sched_getaffinity(0)), 16)
        cpu = cpu_baseline(S, args.events, args.seed + 1, args.cpu_per_worker, workers)

    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    ctx = L.Context(cfg, cond, G, TP.load(cfg.waveform), 0)
    ds = DeviceSteps(S["evkey"], S["counter"], S["edep"], S["z"], S["t"])
    ctx.set_params(cfg, args.seed + 1, ds.evkey_hi)
    n_ev = args.events
    cap_p, cap_c = 90 * n_ev + 1024, 12 * n_ev + 1024
    bufs = {k: torch.empty(cap_p, dtype=torch.int64, device="cuda") for k in ("wf", "seq", "ev")}
    bufs["chan"] = torch.empty(cap_p, dtype=torch.int32, device="cuda")
    for k in L.PULSE_F64:
        bufs[k] = torch.empty(cap_p, dtype=torch.float64, device="cuda")
    for k in L.PULSE_U8:
        bufs[k] = torch.empty(cap_p, dtype=torch.uint8, device="cuda")
    cb = {"cl_event": torch.empty(cap_c, dtype=torch.int64, device="cuda"),
          "cl_mem_off": torch.empty(cap_c + 1, dtype=torch.int64, device="cuda"),
          "mem_idx": torch.empty(cap_p, dtype=torch.int64, device="cuda")}
    for k in ("cl_type", "cl_poscalc"):
        cb[k] = torch.empty(cap_c, dtype=torch.int32, device="cuda")
    cb["cl_sidehits"] = torch.empty(2 * cap_c, dtype=torch.int32, device="cuda")
    for k in L.CLUSTER_F64:
        m = 3 if k == "cl_pos" else 2 if k in ("cl_sidepe", "cl_sidetime") else 1
        cb[k] = torch.empty(m * cap_c, dtype=torch.float64, device="cuda")
    P = L.Pulses(*[L.ptr(bufs[k]) for k in ("wf", "seq", "ev", "chan") + L.PULSE_F64 + L.PULSE_U8], cap_p, 0)
    C = L.Clusters(L.ptr(cb["cl_event"]), L.ptr(cb["cl_type"]), L.ptr(cb["cl_poscalc"]), L.ptr(cb["cl_sidehits"]),
                   *[L.ptr(cb[k]) for k in L.CLUSTER_F64], L.ptr(cb["cl_mem_off"]), cap_c, 0, L.ptr(cb["mem_idx"]),
                   cap_p, 0)
    import ctypes
    st = ds.struct()
    tm = L.Timing()
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ctx.check(ctx.lib.crv_chain_run(ctx.h, ctypes.byref(st), ctypes.byref(P), ctypes.byref(C), ctypes.byref(tm),
                                        ctypes.c_void_p(stream)))
        return {f: getattr(tm, f) for f, _ in L.Timing._fields_}

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runs = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = n_ev * args.steps / dt
    avg = {k: float(np.mean([r[k] for r in runs])) for k in runs[0]}
    cnt = runs[-1]
    S_, W_, P_, C_, M_ = (cnt["n_steps"], cnt["n_records"], cnt["n_pulses"], cnt["n_clusters"], cnt["n_members"])
    b_event = (STEP_B * S_ + RECORD_B * W_ + PULSE_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_) / n_ev
    kernels = {
        "k_sipm": (avg["k_sipm_ms"], STEP_B * S_ + CHARGE_OUT_B * cnt["n_charges"]),
        "k_wf_synth": (avg["k_digi_ms"], CHARGE_IN_B * cnt["n_charges"] + RECORD_B * W_ + DIGI_SLOT_B * cnt["n_digis"]),
        "k_pulses": (avg["k_pulse_ms"], SAMPLE_B * 12 * cnt["n_digis"] + WAVEFORM_META_B * cnt["n_waveforms"]
                     + PULSE_B * P_),
        "k_coinc": (avg["k_coinc_ms"], PULSE_IN_B * P_ + CLUSTER_B * C_ + MEMBER_B * M_),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes = kernels[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload} events, FP64 plain chain, one launch per stage", "events": n_ev,
                   "seed": args.seed, "chain": "steps->photons->SiPM->waveforms->pulses->clusters",
                   "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": kbytes,
                     "ms_per_launch": kms},
        "chain_roofline": {"bytes_per_event": b_event, "achieved": b_event * value / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": b_event * value / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "stage_ms": {k: avg[k] for k in ("charges_ms", "digis_ms", "waveforms_ms", "pulses_ms", "clusters_ms",
                                         "total_ms")},
        "kernel_ms": {k: v[0] for k, v in kernels.items()},
        "per_event": {k[2:]: cnt[k] / n_ev for k in L.TIMING_COUNTS},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
