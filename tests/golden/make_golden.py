"""Write the golden vectors of the reference's numpy CRV stages (run in the development container only).

Imports, read-only, the reference's numpy-only modules from /root/reference/crv/crvfastsim (response/sipm.py,
reco/waveform.py, reco/pulsefit.py, reco/coinc.py, reco/synth.py) and records their outputs on small inputs as .npz
fixtures next to this script. The fixtures hold data only (inputs and outputs); the tests compare the oracle with
them, and nothing that runs the GPU tests reads the reference.

  python tests/golden/make_golden.py            (reference_vectors.npz)
  python tests/golden/make_golden.py noise      (sipm_noise_vectors.npz: dark noise and trap afterpulses)

Pinned here: SiPM charges on a shared uniform stream (the oracle's SequentialSource consumes the reference's draw
order), digis on a shared normal stream, pulse reconstruction, digi grouping, coincidence clusters.
"""
import dataclasses
import os
import sys
import tempfile

sys.dont_write_bytecode = True
os.environ["CRVFS_NO_FETCH"] = "1"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, "/root/reference/crv")
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "mu2e-fastsim_amd"))

import numpy as np  # noqa: E402

from crvfastsim.response import sipm as RSI  # noqa: E402
from crvfastsim.reco import waveform as RWF, pulsefit as RPF, coinc as RC, synth as RSY  # noqa: E402

from crvhip import config as CF, detector as DT, synth as SY, template as TP  # noqa: E402
from oracle import chain as OC  # noqa: E402


def sipm_cases():
    out = {}
    rng = np.random.default_rng(11)
    cases = {
        "a": dict(n=300, mean=60, cfg=dict(cross_talk=0.05), pmap="ones", t_end=1e5),
        "b": dict(n=200, mean=80, cfg=dict(cross_talk=0.5, pixel_id="truncate"), pmap="random", t_end=1e5),
        "c": dict(n=150, mean=40, cfg=dict(cross_talk=0.2), pmap="one_bin", t_end=5000.0),
    }
    for name, c in cases.items():
        n = rng.poisson(c["mean"], c["n"])
        chan = np.repeat(np.arange(c["n"]) * 7 + 3, n)
        t = np.repeat(rng.uniform(0, 1e4, c["n"]), n) + rng.normal(0, 3, len(chan))
        perm = rng.permutation(len(chan))
        chan, t = chan[perm], t[perm]
        photon = np.arange(len(chan)) + 1000
        if c["pmap"] == "ones":
            pmap = np.ones((40, 40))
        elif c["pmap"] == "random":
            pmap = rng.random((40, 40)) ** 4
        else:
            pmap = np.zeros((40, 40))
            pmap[5, 5] = 1.0
        cfg = RSI.SiPMConfig(**c["cfg"])
        seed = 100 + ord(name)
        Q = RSI.simulate(chan, t, photon, pmap, cfg, np.random.default_rng(seed), t_end=c["t_end"])
        pre = f"sipm_{name}_"
        out.update({pre + "chan_in": chan, pre + "time_in": t, pre + "photon_in": photon, pre + "pmap": pmap,
                    pre + "seed": np.int64(seed), pre + "t_end": np.float64(c["t_end"]),
                    pre + "cross_talk": np.float64(cfg.cross_talk), pre + "pixel_id": np.array(cfg.pixel_id)})
        out.update({pre + k: v for k, v in Q.items()})
    return out


def sipm_noise_cases():
    """Reference sipm.simulate with dark noise (thermal_chans over [t_start, t_end)) and trap afterpulses, on fewer
    than its 20000-SiPM chunk so that one chunk holds every SiPM."""
    out = {}
    rng = np.random.default_rng(23)
    cases = {
        "d": dict(n=120, mean=30, cfg=dict(cross_talk=0.05, thermal_rate=0.004), pmap="ones", win=(100.0, 1500.0),
                  extra=40),
        "e": dict(n=150, mean=40, cfg=dict(cross_talk=0.1, trap0_prob=0.08, trap1_prob=0.05), pmap="random",
                  win=(0.0, 5000.0), extra=0),
        "f": dict(n=100, mean=50, cfg=dict(cross_talk=0.3, thermal_rate=0.01, trap0_prob=0.1, trap1_prob=0.1,
                                           trap0_lifetime=7.0, trap1_lifetime=40.0, pixel_id="truncate"),
                  pmap="random", win=(200.0, 1300.0), extra=25),
    }
    for name, c in cases.items():
        n = rng.poisson(c["mean"], c["n"])
        labels = np.arange(c["n"]) * 7 + 3
        chan = np.repeat(labels, n)
        lo, hi = c["win"]
        t = np.repeat(rng.uniform(lo - 50, hi, c["n"]), n) + rng.normal(0, 3, len(chan))
        perm = rng.permutation(len(chan))
        chan, t = chan[perm], t[perm]
        photon = np.arange(len(chan)) + 1000
        pmap = np.ones((40, 40)) if c["pmap"] == "ones" else rng.random((40, 40)) ** 4
        thermal = np.r_[labels, np.arange(c["extra"]) * 7 + 5]
        cfg = RSI.SiPMConfig(**c["cfg"])
        seed = 200 + ord(name)
        Q = RSI.simulate(chan, t, photon, pmap, cfg, np.random.default_rng(seed), t_end=hi, t_start=lo,
                         thermal_chans=thermal)
        pre = f"noise_{name}_"
        out.update({pre + "chan_in": chan, pre + "time_in": t, pre + "photon_in": photon, pre + "pmap": pmap,
                    pre + "thermal": thermal, pre + "seed": np.int64(seed), pre + "t_start": np.float64(lo),
                    pre + "t_end": np.float64(hi), pre + "pixel_id": np.array(cfg.pixel_id)})
        for k in ("cross_talk", "thermal_rate", "trap0_prob", "trap1_prob", "trap0_lifetime", "trap1_lifetime"):
            out[pre + k] = np.float64(getattr(cfg, k))
        out.update({pre + k: v for k, v in Q.items()})
    return out


def chain_inputs(n_events=40, seed=3):
    S = SY.cosmic_events(n_events, seed=seed)
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    Q = OC.charges(S, cond, cfg, seed=5)
    return S, cfg, cond, Q


def waveform_cases(Q, cond):
    out = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "singlePE_synthetic.txt")
        TP.write_raw(path)
        raw = np.loadtxt(path)
        cfg = RWF.WaveformConfig(single_pe_file=path)
        tmpl = RWF.single_pe_template(cfg.single_pe_file, cfg.single_pe_precision, cfg.single_pe_stretch,
                                      cfg.single_pe_max_time)
        out["tmpl_raw"] = raw
        out["tmpl_ref"] = tmpl
        labels = np.unique(Q["chan"])
        tdc0 = np.random.default_rng(9).random(len(labels)) * 12.5
        lut = lambda lab: tdc0[np.searchsorted(labels, lab)]
        out["wf_labels"], out["wf_tdc0"] = labels, tdc0
        out["wf_chan_in"], out["wf_time_in"], out["wf_charge_in"] = Q["chan"], Q["time"], Q["charge"]
        for name, rng in (("quiet", None), ("noise", np.random.default_rng(21))):
            D = RWF.synthesize(Q["chan"], Q["time"], Q["charge"], lut, cfg, rng=rng)
            out.update({f"wf_{name}_{k}": v for k, v in D.items() if not (name == "noise" and k == "volt")})
        # digi grouping on the noisy digis
        D = {k[len("wf_noise_"):]: v for k, v in out.items() if k.startswith("wf_noise_")}
        ev, ch = D["chan"] // CF.CHAN_MULT, D["chan"] % CF.CHAN_MULT
        o = np.lexsort((D["starttdc"], ch, ev))
        lens = np.full(len(o), D["adc"].shape[1], np.int64)
        gid, mat, n, start, first = RPF.concat_waveforms(ev[o], ch[o] // 4, ch[o] % 4, D["starttdc"][o],
                                                         D["adc"][o].ravel(), lens)
        out.update({"cat_gid": gid, "cat_mat": mat, "cat_n": n, "cat_start": start, "cat_first": first})
    return out


def pulse_batches(W_chain):
    b = {}
    adc, n, start, ped, area, height = RSY.waveforms(50, seed=15)
    b["synth12"] = (adc, n, start, ped, area, height)
    adc, n, start, ped, area, height = RSY.waveforms(8, seed=24, L=36)
    adc[0] = ped[0]
    adc[1, :11] = 4095
    adc[2, 0] = ped[2] - 10
    n[3] = 12
    b["edge36"] = (adc, n, start, ped, area, height)
    rng = np.random.default_rng(2)
    L, W = 36, 300
    t = (100 + np.arange(L)) * RPF.PERIOD
    adc = np.zeros((W, L), np.int64)
    for i in range(W):
        y = RPF.gumbel(t, rng.uniform(150, 3000), t[4] + rng.uniform(-6, 6), rng.uniform(12, 24))
        if i % 3 == 0:
            y += RPF.gumbel(t, rng.uniform(100, 1500), t[12] + rng.uniform(-6, 6), rng.uniform(12, 24))
        adc[i] = np.clip(np.round(y + 2045 + rng.normal(0, 2.0, L)), 0, 4095)
    adc[0, :] = 2045
    adc[1, 0] = 2030
    n = np.full(W, L)
    n[5] = 12
    b["double36"] = (adc, n, np.full(W, 100), np.full(W, 2045.0), np.full(W, 400.0), np.full(W, 30.0))
    tt = np.arange(252) * RPF.PERIOD
    sig = RPF.gumbel(tt[None, :], 1000, np.arange(4, 244, 24)[:, None] * RPF.PERIOD, 12).sum(0)
    a = np.rint(sig + 100).astype(np.int64)[None, :].repeat(2, 0)
    a[1, 4] = 4095
    b["train252"] = (a, np.full(2, 252), np.zeros(2, np.int64), np.full(2, 100.), np.full(2, 400.), np.full(2, 11.))
    b["chain"] = tuple(W_chain[k] for k in ("adc", "n", "start", "ped", "area", "height"))
    return b


def pulse_cases(W_chain):
    out = {}
    for name, args in pulse_batches(W_chain).items():
        args = tuple(np.asarray(a) for a in args)
        P = RPF.reconstruct_full(*args)
        for k, v in zip(("adc", "n", "start", "ped", "area", "height"), args):
            out[f"pf_{name}_in_{k}"] = v
        out.update({f"pf_{name}_{k}": v for k, v in P.items()})
    return out


def ref_geom():
    G = DT.sector_geom()
    return RC.SectorGeom(**{f.name: getattr(G, f.name) for f in dataclasses.fields(G)})


def ref_finder():
    c = CF.SYNTH_CUTS
    return RC.FinderConfig(cuts={"CRV": RC.SectorCuts(**dataclasses.asdict(c))})


def coinc_events_from(P, status, G, fcfg):
    """Run the reference finder per event on flat pulses; flatten cluster rows."""
    rows = {k: [] for k in ("cl_event", "cl_type", "cl_pe", "cl_start", "cl_end", "cl_slope", "cl_avgtime",
                            "cl_poscalc", "cl_pos", "cl_sidepe", "cl_sidetime", "cl_sidehits", "cl_nmem")}
    members = []
    ev = P["ev"]
    for e in np.unique(ev):
        a, b = np.searchsorted(ev, e), np.searchsorted(ev, e, side="right")
        p = {k: P[k][a:b] for k in ("idx", "bar", "sipm", "time", "pe", "seq")}
        for r in RC.find_event(p, G, status, fcfg):
            rows["cl_event"].append(e)
            for k, src in (("cl_type", "sectortype"), ("cl_pe", "pe"), ("cl_start", "start"), ("cl_end", "end"),
                           ("cl_slope", "slope"), ("cl_avgtime", "avgtime"), ("cl_poscalc", "poscalc"),
                           ("cl_pos", "pos"), ("cl_sidepe", "sidepe"), ("cl_sidetime", "sidetime"),
                           ("cl_sidehits", "sidehits")):
                rows[k].append(r[src])
            rows["cl_nmem"].append(len(r["pulses"]))
            members += list(r["pulses"])
    out = {k: np.asarray(v) for k, v in rows.items()}
    out["members"] = np.asarray(members, np.int64)
    return out


def coinc_cases(chain_pulses):
    out = {}
    G = ref_geom()
    fcfg = ref_finder()
    status = np.zeros(4 * len(G.sector), np.int64)
    status[np.random.default_rng(1).choice(len(status), 400, replace=False)] = 1
    P = RSY.coinc_events(300, G, seed=7, shower_frac=0.1, shower_mean=40)
    for name, PP, st in (("synth", P, status), ("chain", chain_pulses, np.zeros_like(status))):
        res = coinc_events_from(PP, st, G, fcfg)
        for k in ("ev", "idx", "bar", "sipm", "time", "pe", "seq"):
            out[f"co_{name}_in_{k}"] = np.asarray(PP[k])
        out[f"co_{name}_status"] = st.astype(np.int8)
        out.update({f"co_{name}_{k}": v for k, v in res.items()})
    return out


def main():
    out = {}
    out.update(sipm_cases())
    S, cfg, cond, Q = chain_inputs()
    out.update(waveform_cases(Q, cond))
    # the chain's own reco waveforms (oracle digis; the grouping is pinned separately above)
    tmpl = TP.load(cfg.waveform)
    D = OC.digis(Q, cond, cfg, 5, tmpl)
    W = OC.reco_waveforms(D, cond)
    out.update(pulse_cases(W))
    Pc = OC.pulses(W, cond, cfg)
    wf = Pc["wf"]
    ev = W["ev"][wf]
    o = np.lexsort((Pc["seq"], wf, ev))
    ev, wf = ev[o], wf[o]
    chain_pulses = dict(ev=ev, idx=np.arange(len(ev)) - np.searchsorted(ev, ev), bar=W["bar"][wf],
                        sipm=W["sipm"][wf], time=Pc["time"][o], pe=Pc["pe"][o], seq=Pc["seq"][o])
    out.update(coinc_cases(chain_pulses))
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    if sys.argv[1:] == ["noise"]:
        path = os.path.join(HERE, "sipm_noise_vectors.npz")
        out = sipm_noise_cases()
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path), "bytes,", len(out), "arrays")
    else:
        main()
