"""Synthetic CrvStep tables for the benchmark configurations (host-side input generation, never timed).

Step record (the device input, 32 bytes): event (u32), counter (u32), edep (MeV), z (longitudinal position, mm,
counter centre at 0) and t (ns). Steps come sorted by event. Distributions:

* cosmic: per event Poisson(3) muon tracks; a track picks a module and a counter column uniformly and crosses the
  4 stacked counters there; per step edep ~ Landau(MPV 1.8 MeV, width 0.15 MeV) truncated at 20 MeV, z uniform on
  [-2.55, 2.55] m (one z per track), t uniform on [0, 1695] ns (per track) + 0.3 ns per layer;
* uneven mixture: 90 % cosmic events, 9 % showers of Poisson(40) tracks over 8 adjacent modules (one shower time,
  0.5 ns jitter per track), 1 % dense showers of 200-2,000 steps (log-uniform) in 2 adjacent modules with
  edep ~ Exp(4 MeV), z ~ N(shower z, 300 mm) clipped to the counter, t = shower time + 0.3 ns per layer;
* overlay: any of the above plus Poisson(0.5) background steps per counter, edep ~ Exp(0.8 MeV), z uniform,
  t uniform on [0, 1695] ns.
Randomness is numpy's PCG64 (inputs are data; the chain's own draws are Philox).
"""
import numpy as np

from . import detector as DT

T_RANGE = 1695.0
LAYER_DT = 0.3


class Landau:
    """Inverse-CDF sampler of Landau(mpv, width) truncated at emax, tabulated in double from scipy.stats.landau."""

    def __init__(self, mpv=1.8, width=0.15, emax=20.0, n=200_001):
        from scipy import optimize, stats
        mode = optimize.minimize_scalar(lambda x: -stats.landau.pdf(x), bracket=(-1.0, 0.0, 1.0)).x
        self.mpv, self.width, self.mode = mpv, width, mode
        xmax = mode + (emax - mpv) / width
        x = np.concatenate([np.linspace(-5.0, 20.0, n // 2), np.linspace(20.0, xmax, n - n // 2)[1:]])
        c = stats.landau.cdf(x)
        keep = np.r_[True, np.diff(c) > 0]
        self.x, self.c = x[keep], c[keep]

    def sample(self, rng, size):
        u = rng.uniform(self.c[0], self.c[-1], size)
        return self.mpv + self.width * (np.interp(u, self.c, self.x) - self.mode)


_LANDAU = None


def landau():
    global _LANDAU
    if _LANDAU is None:
        _LANDAU = Landau()
    return _LANDAU


def _tracks(rng, ev_tr, module, col, z, t):
    """4 steps per track (layers 0..3)."""
    T = len(ev_tr)
    layer = np.tile(np.arange(DT.N_LAYERS), T)
    return dict(event=np.repeat(ev_tr, 4), counter=DT.counter_id(np.repeat(module, 4), layer, np.repeat(col, 4)),
                edep=landau().sample(rng, 4 * T), z=np.repeat(z, 4), t=np.repeat(t, 4) + LAYER_DT * layer)


def _cosmic(rng, events):
    ntr = rng.poisson(3.0, len(events))
    ev_tr = np.repeat(events, ntr)
    T = len(ev_tr)
    return _tracks(rng, ev_tr, rng.integers(0, DT.N_MODULES, T), rng.integers(0, DT.N_COLS, T),
                   rng.uniform(-DT.HALF_LENGTH, DT.HALF_LENGTH, T), rng.uniform(0.0, T_RANGE, T))


def _showers(rng, events):
    ntr = rng.poisson(40.0, len(events))
    m0 = rng.integers(0, DT.N_MODULES - 7, len(events))
    t0 = rng.uniform(0.0, T_RANGE, len(events))
    ev_tr = np.repeat(events, ntr)
    T = len(ev_tr)
    return _tracks(rng, ev_tr, np.repeat(m0, ntr) + rng.integers(0, 8, T), rng.integers(0, DT.N_COLS, T),
                   rng.uniform(-DT.HALF_LENGTH, DT.HALF_LENGTH, T), np.repeat(t0, ntr) + rng.normal(0.0, 0.5, T))


def _dense(rng, events):
    n = len(events)
    ns = np.floor(np.exp(rng.uniform(np.log(200.0), np.log(2001.0), n))).astype(np.int64)
    m0 = rng.integers(0, DT.N_MODULES - 1, n)
    z0 = rng.uniform(-DT.HALF_LENGTH, DT.HALF_LENGTH, n)
    t0 = rng.uniform(0.0, T_RANGE, n)
    S = int(ns.sum())
    layer = rng.integers(0, DT.N_LAYERS, S)
    return dict(event=np.repeat(events, ns),
                counter=DT.counter_id(np.repeat(m0, ns) + rng.integers(0, 2, S), layer, rng.integers(0, DT.N_COLS, S)),
                edep=rng.exponential(4.0, S),
                z=np.clip(np.repeat(z0, ns) + rng.normal(0.0, 300.0, S), -DT.HALF_LENGTH, DT.HALF_LENGTH),
                t=np.repeat(t0, ns) + LAYER_DT * layer)


def _background(rng, events, rate=0.5):
    nb = rng.poisson(rate * DT.N_COUNTERS, len(events))
    B = int(nb.sum())
    return dict(event=np.repeat(events, nb), counter=rng.integers(0, DT.N_COUNTERS, B),
                edep=rng.exponential(0.8, B), z=rng.uniform(-DT.HALF_LENGTH, DT.HALF_LENGTH, B),
                t=rng.uniform(0.0, T_RANGE, B))


def _finish(parts):
    parts = [p for p in parts if len(p["event"])]
    if not parts:
        return dict(event=np.zeros(0, np.int64), counter=np.zeros(0, np.int64), edep=np.zeros(0), z=np.zeros(0),
                    t=np.zeros(0), evkey=np.zeros(0, np.int64))
    S = {k: np.concatenate([p[k] for p in parts]) for k in ("event", "counter", "edep", "z", "t")}
    o = np.argsort(S["event"], kind="stable")
    S = {k: v[o] for k, v in S.items()}
    S["event"] = S["event"].astype(np.int64)
    S["counter"] = S["counter"].astype(np.int64)
    S["evkey"] = S["event"].copy()
    return S


def cosmic_events(n_events, seed=0):
    rng = np.random.default_rng(seed)
    return _finish([_cosmic(rng, np.arange(n_events))])


def uneven_events(n_events, seed=0):
    rng = np.random.default_rng(seed)
    kind = rng.choice(3, n_events, p=[0.90, 0.09, 0.01])
    ev = np.arange(n_events)
    return _finish([_cosmic(rng, ev[kind == 0]), _showers(rng, ev[kind == 1]), _dense(rng, ev[kind == 2])])


def overlay_events(n_events, seed=0, rate=0.5, base="uneven"):
    S = uneven_events(n_events, seed) if base == "uneven" else cosmic_events(n_events, seed)
    rng = np.random.default_rng(seed + 0x5EED)
    return _finish([S, _background(rng, np.arange(n_events), rate)])


def step_record(S):
    """The 32-byte device step record as a numpy structured array."""
    rec = np.zeros(len(S["event"]), dtype=[("event", "<u4"), ("counter", "<u4"), ("edep", "<f8"), ("z", "<f8"),
                                            ("t", "<f8")])
    for k in ("event", "counter", "edep", "z", "t"):
        rec[k] = S[k]
    return rec
