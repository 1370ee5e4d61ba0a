"""Opt-in staged onspill response (reference crvfastsim/onspill.py) on the HIP kernels: microbunch-folded photon
times, the 100 ns dark-noise prestart before the digitisation window, dark noise on every live channel of every
event, and the reconstructed bunch correction on pulse times.

The event population (including events without steps) comes in EventTiming; internal event labels are its rows,
returned as event_keys by run(). Steps keep their original times for the optics (the closed-form optics read the
step's time and position); their photons are shifted by their event's mc_pbtime and folded by the period on the
device (crv_photons_run), then go to the SiPM stage with dark noise (crv_sipm_run). Trap afterpulses, fused
responses and empirical backgrounds are refused, as in the reference.
"""
from dataclasses import dataclass, replace
import time as _time

import numpy as np

from . import _lib as L, chain as CH, readout

KEY = ("run", "subrun", "event")


def _ids(keys):
    a = np.asarray(keys)
    if a.ndim != 2 or a.shape[1] != 3 or a.dtype.kind not in "iu" or np.any(a < 0):
        raise ValueError("event keys must be a nonnegative Nx3 integer array (run, subrun, event)")
    return np.array(a, dtype=np.uint64 if a.dtype.kind == "u" else np.int64, copy=True)


def _records(a):
    return np.ascontiguousarray(a).view(np.dtype([(k, a.dtype) for k in KEY])).ravel()


@dataclass(frozen=True)
class EventTiming:
    """Full IDs and per-event times [ns] (reference onspill.py:28-77); S = -mc_pbtime, reco correction = -reco_pbtime."""
    keys: np.ndarray
    mc_pbtime: np.ndarray
    reco_pbtime: np.ndarray
    event_length: np.ndarray
    period: float = 1695.0

    def __post_init__(self):
        keys = _ids(self.keys)
        order = np.argsort(_records(keys), kind="stable")
        ordered = _records(keys[order])
        if np.any(ordered[1:] == ordered[:-1]):
            raise ValueError("event keys must be unique")
        keys.setflags(write=False)
        object.__setattr__(self, "keys", keys)
        object.__setattr__(self, "_order", order)
        object.__setattr__(self, "_ordered", ordered)
        for name in ("mc_pbtime", "reco_pbtime", "event_length"):
            a = np.array(getattr(self, name), dtype=float, copy=True)
            if a.shape != (len(keys),) or not np.isfinite(a).all():
                raise ValueError(f"{name} must be a finite array of length N")
            if name == "event_length" and np.any(a <= 0):
                raise ValueError("event_length must be positive")
            a.setflags(write=False)
            object.__setattr__(self, name, a)
        p = np.asarray(self.period, dtype=float)
        if p.ndim or not np.isfinite(p) or p <= 0:
            raise ValueError("period must be finite and positive")
        object.__setattr__(self, "period", float(p))

    def lookup(self, keys):
        """Dense event rows of full IDs, in caller order (every ID must be in the population)."""
        a = _ids(keys)
        bounds = np.iinfo(self.keys.dtype)
        if a.size and (int(a.min()) < bounds.min or int(a.max()) > bounds.max):
            raise ValueError("step event keys absent from timing")
        rec = _records(a.astype(self.keys.dtype))
        loc = np.searchsorted(self._ordered, rec)
        if np.any(loc >= len(self.keys)) or np.any(self._ordered[np.minimum(loc, len(self.keys) - 1)] != rec):
            raise ValueError("step event keys absent from timing")
        return self._order[loc]


def _nbars(geom):
    n = getattr(geom, "nbars", None)
    return int(n) if n is not None else len(np.asarray(geom.sector))


def _validate(cfg, timing, geom, cond):
    """reference onspill.py:80-107; returns the effective status of the geometry's channels."""
    from .contract import validate_window
    validate_window(cfg.window)
    lo, hi = cfg.window
    if lo < 100 or hi > timing.period or np.any(timing.event_length < hi):
        raise ValueError("onspill window needs a 100 ns prestart and must fit period and every event_length")
    if getattr(cfg, "fused_response", False) or getattr(cfg, "background", None) is not None:
        raise ValueError("onspill does not support fused_response or empirical background")
    s = cfg.sipm
    if s.trap0_prob != 0 or s.trap1_prob != 0:
        raise ValueError("onspill does not support trap afterpulses")
    for name in ("overvoltage", "time_constant", "capacitance", "p1", "p2"):
        if not np.isfinite(getattr(s, name)) or getattr(s, name) <= 0:
            raise ValueError(f"SiPM {name} must be finite and positive")
    if (not np.isfinite(s.thermal_rate) or s.thermal_rate < 0 or s.p1 > 1
            or not np.isfinite(s.cross_talk) or not 0 <= s.cross_talk <= 1):
        raise ValueError("invalid SiPM thermal_rate, p1 or cross_talk")
    for name in ("photon_yield_scale", "photon_yield_variation_scale"):
        if not np.isfinite(getattr(cfg.response, name)) or getattr(cfg.response, name) < 0:
            raise ValueError(f"invalid response {name}")
    if not np.isfinite(cfg.feb_phase_width) or not 0 <= cfg.feb_phase_width <= CH.PERIOD:
        raise ValueError("invalid FEB phase width")
    nchan = _nbars(geom) * 4
    if cond.readout_mask is None:
        raise ValueError("onspill charges require cond.readout_mask; use readout.apply with geometry and counters_cfg")
    status = np.asarray(readout.effective_status(cond))
    if not 0 <= nchan <= CH.CHAN_MULT or status.ndim != 1 or status.dtype.kind not in "iu" or len(status) < nchan:
        raise ValueError("geometry/conditions channels must fit CHAN_MULT and integer status")
    return status[:nchan]


def fold(t, period):
    """Times into [0, period): fmod, then negatives moved up one period (reference onspill.py:91-93)."""
    t = np.fmod(t, period)
    return np.where(t < 0, t + period, t)


def prefilter(ends, shift, period, window):
    """Rows of steps whose folded interval can reach [window[0] - 100, window[1]] (reference onspill.py:128-135):
    a step with finite endpoint times (ends [n, 2]; shift per step) is dropped only when both folded ends lie before
    the prestart, both after the window, or its start after the window and its end before the prestart."""
    lo, hi = window
    finite = np.isfinite(ends).all(axis=1)
    a, b = fold(ends + np.asarray(shift, float)[:, None], period).T
    keep = finite & ~(((a < lo - 100) & (b < lo - 100)) | ((a > hi) & (b > hi)) | ((a > hi) & (b < lo - 100)))
    return np.flatnonzero(keep), finite


_QUIET = {}


def quiet_config(cfg):
    """cfg without its SiPM noise processes: the configuration of the downstream stages and of the device context
    (dark noise reaches the SiPM stage per call)."""
    ent = _QUIET.get(id(cfg))
    if ent is None or ent[0] is not cfg:
        _QUIET.clear()
        ent = (cfg, replace(cfg, sipm=replace(cfg.sipm, thermal_rate=0.0)))
        _QUIET[id(cfg)] = ent
    return ent[1]


def _col(steps, k):
    return np.asarray(steps[k].to_numpy() if hasattr(steps[k], "to_numpy") else steps[k])


def charges(steps, pdg, geom, cond, cfg, rng, *, timing, key=None, tabs=None, pmap=None, timers=None):
    """Steps -> charges sorted by (chan, time); photon is the original input row (-1 dark noise), chan = dense event
    * CHAN_MULT + channel (reference onspill.py:110-187). Every live channel of every event of `timing` gets dark
    noise at cfg.sipm.thermal_rate over [window[0] - 100, window[1]]; photons before that start remain eligible."""
    status = _validate(cfg, timing, geom, cond)
    if tabs is not None:
        raise ValueError("tabs: lookup-table optics are not supported; the closed-form cfg.optics apply")
    cols = [_col(steps, k) for k in KEY]
    bounds = np.iinfo(timing.keys.dtype)
    for col in cols:
        if col.dtype.kind not in "iu":
            raise ValueError("step event keys must have integer columns")
        if col.size and (int(col.min()) < bounds.min or int(col.max()) > bounds.max):
            raise ValueError("step event keys absent from timing")
    ev = timing.lookup(np.column_stack([col.astype(timing.keys.dtype) for col in cols]) if cols[0].size
                       else np.zeros((0, 3), timing.keys.dtype))
    n = len(ev)
    pdg = np.asarray(pdg)
    if pdg.shape != (n,) or (n and pdg.dtype.kind not in "iu"):
        raise ValueError("pdg must be an integer array aligned with every input step")
    ends = np.column_stack([_col(steps, "t0").astype(float), _col(steps, "t1").astype(float)]) if n \
        else np.zeros((0, 2))
    if np.isinf(ends).any():
        raise ValueError("step endpoints must not be infinite")
    lo, hi = cfg.window
    rows, finite = prefilter(ends, timing.mc_pbtime[ev], timing.period, cfg.window)
    timers = timers if timers is not None else {}
    timers["nan_time_steps_skipped"] = timers.get("nan_time_steps_skipped", 0) + int((~finite).sum())
    timers["time_gate_steps_skipped"] = timers.get("time_gate_steps_skipped", 0) + int(finite.sum() - len(rows))
    rows = rows[np.argsort(ev[rows], kind="stable")]
    live = np.flatnonzero((status & 0b101) == 0)
    thermal = (np.arange(len(timing.keys), dtype=np.int64)[:, None] * CH.CHAN_MULT + live).ravel()
    on = cfg.sipm.thermal_rate > 0 and len(thermal) > 0
    if not len(rows) and not on:  # nothing reaches the SiPMs: no device work
        return {k: np.zeros(0, np.int64 if k in ("chan", "photon") else np.float64)
                for k in ("chan", "time", "charge", "pe", "photon")}
    qcfg = quiet_config(cfg)
    CH.CT.validate_chain(qcfg)
    torch = CH._torch()
    ctx = CH.context(qcfg, cond)
    ctx.use_photon_map(qcfg, pmap)
    ctx.set_params(qcfg, CH._seed(key, rng), 0)
    t0 = _time.perf_counter()
    if len(rows):
        sub = {k: _col(steps, k)[rows] for k in _STEP_COLS if k in _keys(steps)}
        counter, edep, z, t = CH.step_columns(sub, geom)
        if np.any(counter < 0) or np.any(counter * 4 + 3 >= len(status)):
            raise ValueError("photon channel outside geometry/conditions")
        ds = CH.DeviceSteps(ev[rows], counter, edep, z, t)
        shift = CH._dev(np.array(timing.mc_pbtime, np.float64), np.float64)
        cap = int(float(np.sum(edep)) * cfg.optics.pe_per_mev * 1.6 * 4 * 0.7) + 4096
        for _ in range(2):
            pb = {k: CH._empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                                     ("photon", torch.int64), ("j", torch.int32),
                                                     ("rank", torch.int32))}
            po = L.Photons(*[L.ptr(pb[k]) for k in ("label", "time", "photon", "j", "rank")], cap, 0, 0)
            rc = ctx.lib.crv_photons_run(ctx.h, CH.C_ref(ds.struct()), L.ptr(shift), float(timing.period),
                                         CH.C_ref(po), None)
            if rc == L.CRV_E_CAPACITY:
                cap = po.n
                continue
            ctx.check(rc)
            break
        nph = po.n
    else:
        pb, nph = None, 0
    timers["photons_s"] = timers.get("photons_s", 0.0) + _time.perf_counter() - t0
    t0 = _time.perf_counter()
    d_th = CH._dev(thermal, np.int64) if on else None
    pl = L.PhotonList(L.ptr(pb["label"]) if nph else None, L.ptr(pb["time"]) if nph else None,
                      L.ptr(pb["photon"]) if nph else None, L.ptr(pb["j"]) if nph else None,
                      L.ptr(pb["rank"]) if nph else None, nph)
    nz = L.SipmNoise(float(cfg.sipm.thermal_rate) if on else 0.0, 0.0, 0.0, 1.0, 1.0, float(lo - 100), float(hi),
                     L.ptr(d_th), len(thermal) if on else 0)
    cap = 2 * nph + 4096
    if on:
        pf = cfg.sipm.p1 * (1.0 - np.exp(-cfg.sipm.overvoltage / cfg.sipm.p2))
        cap += int(3 * cfg.sipm.thermal_rate / pf * (hi - lo + 100) * len(thermal)) + 4096
    for _ in range(2):
        qb = {k: CH._empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                                 ("charge", torch.float64), ("pe", torch.float64),
                                                 ("photon", torch.int64))}
        qo = L.Charges(*[L.ptr(qb[k]) for k in ("label", "time", "charge", "pe", "photon")], cap, 0, 0, 0)
        rc = ctx.lib.crv_sipm_run(ctx.h, CH.C_ref(pl), CH.C_ref(nz), CH.C_ref(qo), None)
        if rc == L.CRV_E_CAPACITY:
            cap = qo.n
            continue
        ctx.check(rc)
        break
    m = qo.n
    photon = CH._host(qb["photon"], m)
    Q = {"chan": CH._host(qb["label"], m), "time": CH._host(qb["time"], m), "charge": CH._host(qb["charge"], m),
         "pe": CH._host(qb["pe"], m), "photon": np.where(photon >= 0, rows[np.maximum(photon, 0)], -1)
         if len(rows) else photon}
    timers["sipm_s"] = timers.get("sipm_s", 0.0) + _time.perf_counter() - t0
    timers["photons"] = timers.get("photons", 0) + int(nph)
    timers["charges"] = timers.get("charges", 0) + m
    return Q


_STEP_COLS = ("counter", "bar", "edep", "z", "z0", "z1", "x0", "y0", "x1", "y1", "t", "t0", "t1")


def _keys(steps):
    return list(steps.keys()) if hasattr(steps, "keys") else list(steps.columns)


def reco_timing(W, timing):
    """Waveform rows carrying event_t0 = -reco_pbtime of their dense event (reference onspill.py:190-195)."""
    ev = np.asarray(W["ev"])
    if ev.ndim != 1 or ev.dtype.kind not in "iu" or np.any(ev < 0) or np.any(ev >= len(timing.keys)):
        raise ValueError("waveform ev must contain dense timing row indices")
    return {**W, "event_t0": -timing.reco_pbtime[ev]}


def run(steps, pdg, geom, cond, cfg, rng, *, timing, counters_cfg, key=None, tabs=None, pmap=None, timers=None,
        detail=False):
    """Staged chain deriving a missing readout mask; event_keys maps dense outputs to original IDs (reference
    onspill.py:198-208)."""
    timers = timers if timers is not None else {}
    if cond.readout_mask is None:
        cond = readout.apply(cond, geom, counters_cfg)
    Q = charges(steps, pdg, geom, cond, cfg, rng, timing=timing, key=key, tabs=tabs, pmap=pmap, timers=timers)
    qcfg = quiet_config(cfg)
    D = CH.digis(Q, cond, qcfg, rng, key=key, timers=timers, stock_scan=True)
    W = reco_timing(CH.reco_waveforms(D, cond), timing)
    clusters, timers = CH.clusters(W, cond, geom, counters_cfg, qcfg, timers=timers, detail=detail, rng=rng)
    return dict(Q=Q, D=D, W=W, clusters=clusters, timers=timers, event_keys=timing.keys)
