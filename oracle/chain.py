"""The CRV response chain in the oracle: steps -> charges -> digis -> reco waveforms -> pulses -> clusters.

Test infrastructure only: tests, the smoke check and the CPU-baseline leg of the benchmark call it as the checker.
Mirrors the reference's crvfastsim/chain.py:151-362 stage functions (charges, tdc0_per_feb + digis,
reco_waveforms, clusters with the numpy pulse reco and the per-event coincidence loop). Configuration objects are
duck-typed (the fields of the reference's SiPMConfig, WaveformConfig, RecoConfig, FinderConfig plus the closed-form
optics); randomness is the Philox layout of oracle.philox keyed by one integer seed.
"""
import numpy as np

from . import coinc as C
from . import pulsefit as PF
from . import response as RS
from . import waveform as WF

CHAN_MULT = 32768


def charges(steps, cond, cfg, seed):
    status = cond.status
    dev = (cond.deviation * cfg.response.photon_yield_variation_scale + 1.0) * cfg.response.photon_yield_scale
    return RS.charges(steps, status, dev, cfg.optics, cfg.sipm, cfg.photon_map, seed, t_end=cfg.window[1])


def digis(Q, cond, cfg, seed, tmpl):
    labels = np.unique(Q["chan"])
    tdc0 = WF.tdc0_per_feb(labels, cond, seed, width=cfg.feb_phase_width, key=cfg.feb_phase_key)
    lut = lambda lab: tdc0[np.searchsorted(labels, lab)]
    noise = WF.PhiloxNoise(seed) if cfg.noise else None
    return WF.synthesize(Q["chan"], Q["time"], Q["charge"], lut, tmpl, cfg.waveform, window=cfg.window,
                         noise=noise, window_samples=cfg.window_samples)


def reco_waveforms(D, cond):
    return WF.reco_waveforms(D, cond)


def pulses(W, cond, cfg):
    rcfg = cfg.pulse_reco
    P = PF.reconstruct_full(W["adc"], W["n"], W["start"], W["ped"], W["area"], W["height"], rcfg)
    P = PF.filter_channels(P, W['bar'] * 4 + W['sipm'], cond.status, rcfg)
    if rcfg.use_time_offset_db:
        P["time"] = P["time"] + W["toff"][P["wf"]] * rcfg.time_offset_scale
    return P


def clusters(W, cond, G, cfg, P=None):
    """Reference chain.py:296-362 numpy branch, detail=True. Returns (out dict, P)."""
    if P is None:
        P = pulses(W, cond, cfg)
    wf = P["wf"]
    ev = W["ev"][wf]
    o = np.lexsort((P["seq"], wf, ev))
    ev, wf = ev[o], wf[o]
    idx = np.arange(len(ev)) - np.searchsorted(ev, ev)
    bar, sipm, time, pe, seq = W["bar"][wf], W["sipm"][wf], P["time"][o], P["pe"][o], P["seq"][o]
    out = find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    out["pulses"] = dict(ev=ev, idx=idx, bar=bar, sipm=sipm, seq=seq, pe=pe, time=time)
    return out, P


def find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G):
    """Pulses sorted by event (collection order within an event) -> cluster table: coinc.find_event per event
    (reference chain.py:337-362)."""
    ev = np.asarray(ev)
    idx = np.arange(len(ev)) - np.searchsorted(ev, ev)
    cl = {k: [] for k in ("cl_event", "cl_type", "cl_pe", "cl_start", "cl_end", "cl_slope", "cl_avgtime", "cl_poscalc")}
    mem = {"cl_pos": [], "mem_cl": [], "mem_idx": [], "mem_event": []}
    side = {"cl_sidehits": [], "cl_sidepe": [], "cl_sidetime": []}
    ue, lo = np.unique(ev, return_index=True)
    hi = np.r_[lo[1:], len(ev)]
    for e, a, b in zip(ue, lo, hi):
        p = dict(idx=idx[a:b], bar=bar[a:b], sipm=sipm[a:b], time=time[a:b], pe=pe[a:b], seq=seq[a:b])
        for r in C.find_event(p, G, cond.status, cfg.finder):
            cl["cl_event"].append(e)
            for k, src in (("cl_type", "sectortype"), ("cl_pe", "pe"), ("cl_start", "start"), ("cl_end", "end"),
                           ("cl_slope", "slope"), ("cl_avgtime", "avgtime"), ("cl_poscalc", "poscalc")):
                cl[k].append(r[src])
            for k, src in (("cl_sidehits", "sidehits"), ("cl_sidepe", "sidepe"), ("cl_sidetime", "sidetime")):
                side[k].append(np.asarray(r[src]))
            n = len(cl["cl_event"]) - 1
            mem["cl_pos"].append(np.asarray(r["pos"], float))
            mem["mem_cl"] += [n] * len(r["pulses"])
            mem["mem_idx"] += list(r["pulses"])
            mem["mem_event"] += [e] * len(r["pulses"])
    out = {k: np.asarray(v) for k, v in cl.items()}
    out.update({k: np.asarray(v, np.int64) for k, v in mem.items() if k != "cl_pos"})
    out["cl_pos"] = np.asarray(mem["cl_pos"], float).reshape(-1, 3)
    out["cl_sidehits"] = np.asarray(side["cl_sidehits"], np.int64).reshape(-1, 2)
    out["cl_sidepe"] = np.asarray(side["cl_sidepe"], float).reshape(-1, 2)
    out["cl_sidetime"] = np.asarray(side["cl_sidetime"], float).reshape(-1, 2)
    return out


def run(steps, cond, G, cfg, seed, tmpl):
    """Whole chain; returns dict of stage outputs."""
    Q = charges(steps, cond, cfg, seed)
    D = digis(Q, cond, cfg, seed, tmpl)
    W = reco_waveforms(D, cond)
    out, P = clusters(W, cond, G, cfg)
    return {"Q": Q, "D": D, "W": W, "P": P, "clusters": out}
