"""CRV waveform / digitisation oracle (test infrastructure only).

Restates the reference's
  reco/waveform.py:46-62 (single_pe_template: resampling of the 1PE template text file),
  reco/waveform.py:65-71 (charge_clusters: split at gaps > singlePEWaveformMaxTime + 4 periods),
  reco/waveform.py:96-172 (_synthesize: first FEB tick, template sums, noise, zero suppression, ADC),
  chain.py:219-231 (tdc0_per_feb), chain.py:273-286 + reco/pulsefit.py:319-345 (reco_waveforms grouping).

Two spec additions of the synthetic build, both off when `window_samples` is None (then this is the reference):
* a charge cluster's waveform is a fixed window of `window_samples` (128) samples: template contributions and
  zero-suppression starts are limited to the first window_samples - samples_zs samples (clusters that would be
  longer are counted as 'overlong');
* ADC conversion saturates: clip(trunc(clip(x, -1, max_adc + 1)), min_adc, max_adc), equal to the reference's
  int16 cast + clip wherever that cast is defined.
Noise: the reference's rng.normal over the [C, L] array (SequentialSource) or Box-Muller on Philox uniforms keyed by
(event key, channel, window tick, sample // 2) (PhiloxNoise).
"""
import numpy as np

from . import philox as PX
from . import rand as R

PERIOD = 12.5
N_DIGI_PERIODS_GAP = 4
CHAN_MULT = 32768


def single_pe_template(raw_t, raw_v, precision, stretch, max_time):
    """Reference waveform.py:47-62 on the two columns of the template file."""
    t, v = np.asarray(raw_t, float) * stretch, np.asarray(raw_v, float)
    out = [v[0]]
    idx = 1
    for i in range(1, len(t)):
        tt = idx * precision
        while t[i] >= tt and idx * precision < max_time:
            frac = (tt - t[i - 1]) / (t[i] - t[i - 1])
            out.append((v[i] - v[i - 1]) * frac + v[i - 1])
            idx += 1
            tt = idx * precision
        if idx * precision >= max_time:
            break
    return np.asarray(out)


def charge_clusters(chan, time, cfg):
    gap = cfg.single_pe_max_time + N_DIGI_PERIODS_GAP * PERIOD
    new = np.ones(len(time), bool)
    if len(time) > 1:
        new[1:] = (chan[1:] != chan[:-1]) | (time[1:] - time[:-1] > gap)
    return np.cumsum(new) - 1


def tdc0_per_feb(labels, cond, seed, width=PERIOD, key="roc_feb", chan_mult=CHAN_MULT):
    """One uniform x width per (event, FEB key); FEB key = roc * 1000 + feb ("roc_feb") or feb ("feb")."""
    labels = np.asarray(labels, np.int64)
    ev, ch = labels // chan_mult, labels % chan_mult
    fk = cond.feb[ch] if key == "feb" else cond.roc[ch] * 1000 + cond.feb[ch]
    u = PX.draw(seed, PX.KIND_TDC0, ev, fk, 0, 0)[0]
    return u * width


class PhiloxNoise:
    def __init__(self, seed, chan_mult=CHAN_MULT):
        self.seed, self.chan_mult = seed, chan_mult

    def __call__(self, clabel, tick, L, valid, sigma):
        C = len(clabel)
        out = np.zeros((C, L))
        if C == 0 or L == 0:
            return out
        ev = np.repeat(clabel // self.chan_mult, L)
        ch = np.repeat(clabel % self.chan_mult, L)
        tk = np.repeat(tick, L)
        J = np.tile(np.arange(L), C)
        u = PX.draw(self.seed, PX.KIND_NOISE, ev, ch, tk.astype(np.int64) & 0xFFFFFFFF, J >> 1)
        odd = (J & 1) == 1
        z = R.normal(np.where(odd, u[2], u[0]), np.where(odd, u[3], u[1]))
        return np.where(valid, sigma * z.reshape(C, L), 0.0)


class SequentialNoise:
    def __init__(self, rng):
        self.rng = rng

    def __call__(self, clabel, tick, L, valid, sigma):
        return self.rng.normal(0.0, sigma, size=(len(clabel), L)) * valid


def adc_of(volt, cfg):
    x = (volt * cfg.adc_conversion + cfg.pedestal) + 0.5
    return np.clip(np.trunc(np.clip(x, -1.0, cfg.max_adc + 1.0)), cfg.min_adc, cfg.max_adc).astype(np.int64)


def synthesize(chan, time, charge, tdc0, tmpl, cfg, window=(0.0, 1e5), noise=None, window_samples=128):
    """Zero-suppressed digis (reference reco/waveform.py:96-172). chan: labels grouped and time-ordered; tdc0: per
    charge-cluster lookup callable(labels) or array by label. noise: PhiloxNoise / SequentialNoise / None.

    Returns dict chan, start, tdc0, starttdc, volt [D, S], adc [D, S] (S = samples_zs) sorted by (chan, start), the
    per-cluster tables (cluster_chan, cluster_tstart, cluster_tick, cluster_keep, cluster_length) and, with
    window_samples, 'record_adc' [C, window_samples] of every cluster that has a digi ('record_cluster').
    """
    chan = np.asarray(chan, np.int64)
    time = np.asarray(time, np.float64)
    charge = np.asarray(charge, np.float64)
    tmpl = np.asarray(tmpl, np.float64)
    Sz = cfg.samples_zs
    vmax = tmpl.max()
    cid = charge_clusters(chan, time, cfg)
    C = int(cid[-1]) + 1 if len(cid) else 0
    first = np.flatnonzero(np.r_[True, cid[1:] != cid[:-1]]) if C else np.zeros(0, np.int64)
    last = np.r_[first[1:] - 1, len(cid) - 1] if C else np.zeros(0, np.int64)
    nq = np.bincount(cid, minlength=C)
    keep = nq * vmax >= cfg.min_voltage
    cchan = chan[first]
    td = tdc0(cchan) if callable(tdc0) else np.asarray(tdc0)[cchan]
    t_first = time[first] - PERIOD
    kf = np.ceil((t_first - td) / PERIOD)
    t_start = kf * PERIOD + td
    est = np.floor((time[last] - time[first] + cfg.single_pe_max_time) / PERIOD).astype(np.int64)
    nreach = int(np.ceil(cfg.single_pe_max_time / PERIOD)) + 2
    q_ts = t_start[cid]
    j0 = np.maximum(np.ceil((time - q_ts) / PERIOD), 0).astype(np.int64)
    J = j0[:, None] + np.arange(nreach)[None, :]
    st = J * PERIOD + q_ts[:, None] - time[:, None]
    ti = np.rint(st / cfg.single_pe_precision).astype(np.int64)
    ok = (st >= 0) & (ti < len(tmpl)) & keep[cid][:, None]
    lmax = None if window_samples is None else window_samples - Sz
    used_full = np.zeros(C, np.int64)
    if ok.any():
        np.maximum.at(used_full, np.repeat(cid, nreach)[ok.ravel()], J.ravel()[ok.ravel()] + 1)
    if lmax is not None:
        ok &= J < lmax
    used_len = np.zeros(C, np.int64)
    if ok.any():
        np.maximum.at(used_len, np.repeat(cid, nreach)[ok.ravel()], J.ravel()[ok.ravel()] + 1)
    length = np.maximum(est, used_len)
    overlong = np.zeros(C, bool)
    if lmax is not None:
        overlong = (np.maximum(est, used_full) > lmax) & keep
        length = np.minimum(length, lmax)
    L = int(length.max()) + Sz if C else 0
    if window_samples is not None:
        L = max(L, window_samples)
    wave = np.zeros((C, L))
    vals = tmpl[np.clip(ti, 0, len(tmpl) - 1)] * (charge / cfg.single_pe_reference_charge)[:, None]
    np.add.at(wave, (np.repeat(cid, nreach)[ok.ravel()], J.ravel()[ok.ravel()]), vals.ravel()[ok.ravel()])
    valid = np.arange(L)[None, :] < length[:, None]
    if noise is not None:
        wave += noise(cchan, kf.astype(np.int64), L, valid, cfg.noise)
    above = (wave > cfg.min_voltage) & valid
    win = np.zeros_like(above)
    for d in range(-2, 3):
        sh = np.roll(above, -d, axis=1)
        if d > 0:
            sh[:, L - d:] = False
        elif d < 0:
            sh[:, :-d] = False
        win |= sh
    win &= valid & keep[:, None]
    next_i = np.zeros(C, np.int64)
    starts_r, starts_i = [], []
    for i in range(L):
        cand = win[:, i] & (next_i <= i)
        if not cand.any():
            continue
        tstart = t_start + i * PERIOD
        inwin = cand & (tstart >= window[0]) & (tstart <= window[1])
        r = np.flatnonzero(inwin)
        starts_r.append(r)
        starts_i.append(np.full(len(r), i))
        next_i[r] = i + Sz
    if starts_r:
        r = np.concatenate(starts_r)
        i0 = np.concatenate(starts_i)
    else:
        r = i0 = np.zeros(0, np.int64)
    cols = i0[:, None] + np.arange(Sz)[None, :]
    volt = np.where(cols < length[r, None], wave[r[:, None], np.minimum(cols, L - 1)], 0.0)
    start = t_start[r] + i0 * PERIOD
    adc = adc_of(volt, cfg)
    order = np.lexsort((start, cchan[r]))
    out = {"chan": cchan[r][order], "start": start[order], "tdc0": td[r][order],
           "starttdc": np.rint((start - td[r]) / PERIOD).astype(np.int64)[order],
           "volt": volt[order], "adc": adc[order], "digi_cluster": r[order], "digi_sample": i0[order],
           "cluster_chan": cchan, "cluster_tstart": t_start, "cluster_tick": kf.astype(np.int64),
           "cluster_keep": keep, "cluster_length": length, "overlong": int(overlong.sum())}
    if window_samples is not None:
        rc = np.unique(r)
        rw = np.where(np.arange(window_samples)[None, :] < length[rc, None], wave[rc, :window_samples], 0.0)
        out["record_cluster"] = rc
        out["record_adc"] = adc_of(rw, cfg)
    return out


def concat_waveforms(key, bar, sipm, start_tdc, adc_flat, lens):
    """Reference reco/pulsefit.py:319-345."""
    nd = len(bar)
    lens = np.asarray(lens, dtype=np.int64)
    new = np.ones(nd, bool)
    if nd > 1:
        new[1:] = ~((key[1:] == key[:-1]) & (bar[1:] == bar[:-1]) & (sipm[1:] == sipm[:-1]) &
                    (start_tdc[1:] == start_tdc[:-1] + lens[:-1]))
    gid = np.cumsum(new) - 1
    G = int(gid[-1]) + 1 if nd else 0
    glen = np.bincount(gid, weights=lens, minlength=G).astype(np.int64)
    L = int(glen.max()) if G else 0
    first = np.flatnonzero(new)
    offset = np.zeros(nd, np.int64)
    if nd:
        csum = np.cumsum(lens) - lens
        offset = csum - csum[first][gid]
    mat = np.zeros((G, L), np.int64)
    within = np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens)
    col = np.repeat(offset, lens) + within
    mat[np.repeat(gid, lens), col] = np.asarray(adc_flat, dtype=np.int64)
    return gid, mat, glen, start_tdc[first], first


def reco_waveforms(D, cond, chan_mult=CHAN_MULT):
    """Reference chain.py:273-286."""
    ev, ch = D["chan"] // chan_mult, D["chan"] % chan_mult
    o = np.lexsort((D["starttdc"], ch, ev))
    ev, ch, st = ev[o], ch[o], D["starttdc"][o]
    adc = np.asarray(D["adc"])[o]
    lens = np.full(len(o), adc.shape[1] if adc.ndim == 2 else 0, np.int64)
    gid, mat, n, start, first = concat_waveforms(ev, ch // 4, ch % 4, st, adc.ravel(), lens)
    c0 = ch[first]
    return {"ev": ev[first], "bar": c0 // 4, "sipm": c0 % 4, "chan": c0, "adc": mat, "n": n, "start": start,
            "ped": cond.pedestal[c0], "area": cond.pulsearea[c0], "height": cond.pulseheight[c0],
            "toff": cond.timeoffset[c0]}
