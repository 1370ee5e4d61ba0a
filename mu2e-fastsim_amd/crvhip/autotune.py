"""Work-queue autotuner of the FP32 batched chain.

The SiPM and waveform kernels of crv_chain_run_batched are persistent grids that take tiles of SiPMs from an atomic
work queue, optionally longest first (by photon count, or by its octave or quarter octave, which keeps SiPMs that
share steps together). The best tile and order depend on the workload's spread of photon counts (cosmic events: tens
of photons per SiPM; dense showers: thousands on a few SiPMs) and on the device, so they are measured:
every candidate (tile, sort order, grid) runs the whole chain on a sample of the workload with zero-capacity outputs
(the library then runs every kernel, reports its timing and writes nothing), and the fastest by the median of the
work-queue kernels' times is kept. Results do not depend on the choice (the draws are keyed by event, channel, step
and photon, not by lane or tile), so tuning changes speed only.

Choices are cached per (device, precision, workload signature) in a JSON file: $CRVHIP_TUNE_CACHE, else
~/.cache/crvhip/autotune.json.
"""
import ctypes
import json
import math
import os

from . import _lib as L

TILES = (64, 128, 256, 512)
SORTS = (1, 3, 2, 0)  # crv_batch_opts.sort_tasks: photon count, its quarter octave, its octave, as given
GRIDS = (0,)


def cache_path():
    return os.environ.get("CRVHIP_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "crvhip",
                                                               "autotune.json")


def signature(device_name, n_events, n_steps, precision="fp32"):
    """Cache key: device, precision and the workload's steps per event on a log2 scale (cosmic ~12, uneven ~33,
    overlay ~2,800 fall in different buckets)."""
    spe = n_steps / max(n_events, 1)
    return f"{device_name}|{precision}|spe2^{int(math.floor(math.log2(max(spe, 1.0))))}"


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store_cache(entries, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def pick(results):
    """The candidate with the smallest median work-queue time; ties go to the earlier candidate.
    results: list of (opts dict, [times in ms])."""
    best, best_t = None, math.inf
    for opts, times in results:
        ts = sorted(times)
        med = ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2])
        if med < best_t:
            best, best_t = opts, med
    return best, best_t


def _empty_outputs():
    P = L.Pulses()
    C = L.Clusters()
    P.capacity = C.capacity = C.mem_capacity = 0
    return P, C


def measure(ctx, steps, opts, repeats=3, stream=None):
    """Work-queue kernel times (ms: SiPM + waveform + fit kernels) of `repeats` zero-capacity runs of the batched
    chain over the device steps `steps` (a crv_steps), after one untimed run with the same options; infinite for a
    candidate whose workspace does not fit in device memory."""
    P, C = _empty_outputs()
    tm = L.Timing()
    bo = L.BatchOpts(**opts)
    times = []
    for i in range(repeats + 1):
        rc = ctx.lib.crv_chain_run_batched(ctx.h, ctypes.byref(steps), ctypes.byref(bo), ctypes.byref(P),
                                           ctypes.byref(C), ctypes.byref(tm), stream)
        if rc not in (0, L.CRV_E_CAPACITY):
            try:
                ctx.check(rc)
            except L.CrvError as e:
                if "workspace allocation failed" in str(e):  # the candidate does not fit the device: never chosen
                    return [math.inf] * repeats
                raise
        if i:
            times.append(tm.k_sipm_ms + tm.k_digi_ms + tm.k_pulse_ms)
    return times


def tune(ctx, steps, n_events, device_name="", tiles=TILES, sorts=SORTS, grids=GRIDS, repeats=3, stream=None,
         cache=True, cache_file=None):
    """(grid, tile, sort_tasks) of the fastest candidate on `steps` (a crv_steps over n_events events resident on the
    device; ctx a crvhip._lib.Context with parameters set). Returns (opts dict, report dict)."""
    key = signature(device_name, n_events, steps.n)
    entries = load_cache(cache_file) if cache else {}
    if cache and key in entries:
        e = entries[key]
        return dict(e["opts"]), {"cached": True, "key": key, "median_ms": e["median_ms"]}
    results = []
    for g in grids:
        for s in sorts:
            for t in tiles:
                opts = {"grid": g, "tile": t, "sort_tasks": s}
                results.append((opts, measure(ctx, steps, opts, repeats, stream)))
    best, best_t = pick(results)
    report = {"cached": False, "key": key, "median_ms": best_t,
              "candidates": [{"opts": o, "ms": t} for o, t in results]}
    if cache:
        entries[key] = {"opts": best, "median_ms": best_t, "events": n_events}
        store_cache(entries, cache_file)
    return dict(best), report
