"""Out-of-time CRV background pulses added before the coincidence finder (reference reco/background.py:112-212).

A BackgroundModel holds a library of pulse groups (one event's out-of-time pulses of one sector slot, chained in
time) and a net rate. sample() draws, per event, a group count (Poisson, optionally gamma-mixed for the data's
overdispersion), groups uniformly from the library and a uniform start time in the window; merge_pulses() adds them
to the signal pulses in (event, bar, SiPM, time) order. The library is data-derived and comes from the caller (save /
load as npz); fitting it needs the experiment's ntuples (uproot / awkward), which this package does not read.
Sampling is host work on a numpy Generator, in the reference's draw order; the finder that follows runs on the GPU.
"""
from dataclasses import dataclass, field
import json

import numpy as np

LIB_KEYS = ("group", "bar", "sipm", "pe", "dt", "seq")


@dataclass
class BackgroundModel:
    rate_per_ns: float
    gamma_k: float
    start: np.ndarray
    count: np.ndarray
    bar: np.ndarray
    sipm: np.ndarray
    pe: np.ndarray
    dt: np.ndarray
    seq: np.ndarray
    info: dict = field(default_factory=dict)

    @classmethod
    def from_partials(cls, data_parts, mc_parts=()):
        """The model from per-file partial fits (reference background.py:129-166): net rate = data groups minus MC
        groups per event-ns (never below 0); gamma shape of the data's group count from its mean and variance
        (infinite without overdispersion); library = the kept groups of every data file, renumbered file by file.
        A part holds events, exposure_ns, groups, cnt_sum, cnt_sumsq, lib (LIB_KEYS arrays), file and keep_frac."""
        total = lambda parts, key: sum(part[key] for part in parts)
        events, exposure = total(data_parts, "events"), total(data_parts, "exposure_ns")
        groups = total(data_parts, "groups")
        data_rate = groups / exposure if exposure else 0.0
        mc_exposure = total(mc_parts, "exposure_ns") if mc_parts else 0
        mc_rate = total(mc_parts, "groups") / mc_exposure if mc_exposure else 0.0
        mean = total(data_parts, "cnt_sum") / max(events, 1)
        var = total(data_parts, "cnt_sumsq") / max(events, 1) - mean * mean
        shape = float(mean * mean / (var - mean)) if var > mean else np.inf
        cols, base = {key: [] for key in LIB_KEYS}, 0
        for part in data_parts:
            grp = np.asarray(part["lib"]["group"], np.int64)
            cols["group"].append(grp + base)
            for key in LIB_KEYS[1:]:
                cols[key].append(np.asarray(part["lib"][key]))
            base += int(grp.max()) + 1 if grp.size else 0
        lib = {key: (np.concatenate(v) if v else np.zeros(0)) for key, v in cols.items()}
        by_group = np.argsort(lib["group"], kind="stable")
        lib = {key: v[by_group] for key, v in lib.items()}
        n_lib = int(lib["group"].max()) + 1 if lib["group"].size else 0
        count = np.bincount(lib["group"].astype(np.int64), minlength=n_lib)
        info = {"data_window_ns": float(exposure / max(events, 1)), "data_events": int(events),
                "data_groups": int(groups), "library_groups": n_lib,
                "data_files": [part.get("file") for part in data_parts],
                "mc_files": [part.get("file") for part in mc_parts],
                "keep_frac": sorted({part.get("keep_frac") for part in data_parts}),
                "data_rate_per_event_per_us": 1e3 * data_rate, "mc_rate_per_event_per_us": 1e3 * mc_rate,
                "pulses_per_group": float(lib["bar"].size / max(n_lib, 1)), "gamma_k_window": shape}
        return cls(rate_per_ns=max(data_rate - mc_rate, 0.0), gamma_k=shape,
                   start=(np.cumsum(count) - count).astype(np.int64), count=count.astype(np.int64),
                   bar=lib["bar"].astype(np.int64), sipm=lib["sipm"].astype(np.int64), pe=lib["pe"].astype(float),
                   dt=lib["dt"].astype(float), seq=lib["seq"].astype(np.int64), info=info)

    def sample(self, events, window, rng, rate_scale=1.0, overdispersed=True):
        """Background pulses of the event keys `events` in the window (w0, w1) [ns]: dict ev, bar, sipm, pe, time, seq
        (flat, unsorted). Draws from `rng` in the reference's order (background.py:168-193): the per-event gamma
        means (their shape scaled from the data window to this one), the Poisson group counts, the library groups,
        then their start times."""
        keys = np.asarray(events, np.int64)
        w0, w1 = window
        if not keys.size or rate_scale <= 0 or self.rate_per_ns <= 0 or not len(self.count):
            return _no_pulses()
        mean = self.rate_per_ns * rate_scale * (w1 - w0)
        if overdispersed and np.isfinite(self.gamma_k):
            shape = self.gamma_k * (w1 - w0) / self.info.get("data_window_ns", w1 - w0)
            lam = rng.gamma(shape, mean / shape, keys.size)
        else:
            lam = np.full(keys.size, mean)
        per_event = rng.poisson(lam)
        n_groups = int(per_event.sum())
        if n_groups == 0:
            return _no_pulses()
        which = rng.integers(0, len(self.count), n_groups)
        t_start = w0 + (w1 - w0) * rng.random(n_groups)
        sizes = self.count[which]
        member = np.repeat(np.arange(n_groups), sizes)          # group of each sampled pulse
        first = np.cumsum(sizes) - sizes
        row = self.start[which][member] + (np.arange(member.size) - first[member])
        return {"ev": np.repeat(keys, per_event)[member], "bar": self.bar[row], "sipm": self.sipm[row],
                "pe": self.pe[row], "time": t_start[member] + self.dt[row], "seq": self.seq[row]}

    def save(self, path):
        np.savez(path, rate_per_ns=self.rate_per_ns, gamma_k=self.gamma_k, start=self.start, count=self.count,
                 bar=self.bar, sipm=self.sipm, pe=self.pe, dt=self.dt, seq=self.seq, info=json.dumps(self.info))

    @classmethod
    def load(cls, path):
        """A library written by save() (or the reference's BackgroundModel.save): plain arrays, no pickles."""
        z = np.load(path, allow_pickle=False)
        return cls(rate_per_ns=float(z["rate_per_ns"]), gamma_k=float(z["gamma_k"]), start=z["start"],
                   count=z["count"], bar=z["bar"], sipm=z["sipm"], pe=z["pe"], dt=z["dt"], seq=z["seq"],
                   info=json.loads(str(z["info"])))


def _no_pulses():
    return {"ev": np.zeros(0, np.int64), "bar": np.zeros(0, np.int64), "sipm": np.zeros(0, np.int64),
            "pe": np.zeros(0), "time": np.zeros(0), "seq": np.zeros(0, np.int64)}


def merge_pulses(ev, idx, bar, sipm, tm, pe, seq, B):
    """Signal pulses (sorted by event) plus background pulses B -> one table sorted by (event, bar, SiPM, time), idx
    recomputed per event, and a flag per pulse (True: background) (reference background.py:199-212)."""
    cols = [np.concatenate([np.asarray(a), np.asarray(B[k])]) for a, k in
            ((ev, "ev"), (bar, "bar"), (sipm, "sipm"), (tm, "time"), (pe, "pe"), (seq, "seq"))]
    bg = np.concatenate([np.zeros(len(ev), bool), np.ones(len(B["ev"]), bool)])
    order = np.lexsort((cols[3], cols[2], cols[1], cols[0]))
    e, b, s, t, p, q = (c[order] for c in cols)
    return e, np.arange(e.size) - np.searchsorted(e, e), b, s, t, p, q, bg[order]


def sampler(cfg, rng):
    """Callable(events) -> background pulses, or None without a model (reference chain.py:288-293)."""
    if getattr(cfg, "background", None) is None or cfg.background_scale <= 0:
        return None
    return lambda events: cfg.background.sample(events, cfg.window, rng, rate_scale=cfg.background_scale)
