"""Onspill staged response oracle (test infrastructure only): reference crvfastsim/onspill.py:101-187 under the
closed-form optics.

  * a step is kept unless both endpoint times, shifted by its event's S = mc_pbtime and folded into [0, period), lie
    before lo - 100, or both after hi, or its start after hi and its end before lo - 100 (onspill.py:128-135);
  * the kept steps' photons (their original times; oracle.response.photons on the dense event rows) are shifted by S
    and folded, and labelled dense event * CHAN_MULT + channel (onspill.py:141-150);
  * every live channel of every event gets dark noise over [lo - 100, hi] and the SiPM simulation keeps photons up
    to hi (onspill.py:152-161).
The photons go to the SiPM stage in (label, step ordinal, rank) order with their (dense event, channel, step ordinal,
rank) draw identity, as the HIP path hands them over.
"""
import numpy as np

from . import response as RS

CHAN_MULT = RS.CHAN_MULT


def fold(t, period):
    """reference onspill.py:91-93: fmod into (-period, period), then negatives moved up by one period."""
    t = np.fmod(t, period)
    return np.where(t < 0, t + period, t)


def gate(t0, t1, shift, period, lo, hi):
    """Rows of the steps the prefilter keeps (reference onspill.py:128-135); shift per step."""
    ends = np.column_stack([np.asarray(t0, float), np.asarray(t1, float)])
    finite = np.isfinite(ends).all(axis=1)
    a, b = fold(ends + np.asarray(shift, float)[:, None], period).T
    return np.flatnonzero(finite & ~(((a < lo - 100) & (b < lo - 100)) | ((a > hi) & (b > hi))
                                     | ((a > hi) & (b < lo - 100))))


def charges(ev, counter, edep, z, t, t0, t1, status, dev, opt, sipm_cfg, pmap, seed, mc_pbtime, period, window,
            n_events, nchan):
    """Steps (dense event row ev, counter, edep, z, photon time t, endpoint times t0 / t1; arrays over the input
    rows) -> SiPM charges dict (chan label, time, charge, pe, photon = input row, -1 dark), sorted by (chan, time)."""
    lo, hi = window
    ev = np.asarray(ev, np.int64)
    rows = gate(t0, t1, np.asarray(mc_pbtime, float)[ev], period, lo, hi)
    rows = rows[np.argsort(ev[rows], kind="stable")]  # the step table of the device call, sorted by event
    S = {"evkey": ev[rows], "counter": np.asarray(counter)[rows], "edep": np.asarray(edep, float)[rows],
         "z": np.asarray(z, float)[rows], "t": np.asarray(t, float)[rows]}
    if len(rows):
        P = RS.photons(S, status, dev, opt, seed)
    else:
        P = {k: np.zeros(0, np.float64 if k == "time" else np.int64) for k in ("evkey", "chan", "row", "j", "rank",
                                                                             "time")}
    label = P["evkey"] * CHAN_MULT + P["chan"]
    o = np.lexsort((P["rank"], P["j"], label))
    P = {k: (v[o] if isinstance(v, np.ndarray) and len(v) == len(label) else v) for k, v in P.items()}
    label = label[o]
    time = fold(P["time"] + np.asarray(mc_pbtime, float)[P["evkey"]], period)
    live = np.flatnonzero((np.asarray(status)[:nchan] & 0b101) == 0)
    thermal = (np.arange(n_events, dtype=np.int64)[:, None] * CHAN_MULT + live).ravel()
    src = RS.PhiloxSource(seed, (P["evkey"], P["chan"], P["j"], P["rank"]))
    Q = RS.sipm_simulate(label, time, rows[P["row"]] if len(label) else np.zeros(0, np.int64), pmap, sipm_cfg, src,
                         t_end=hi, t_start=lo - 100, thermal_labels=thermal)
    return Q
