"""Lookup-table optics on the GPU (crv_lookup_photons_run through the C-ABI) against the oracle's restatement of the
reference's response/photons.py, bit for bit.

Both sides draw the photon numbers, the Binomial arrivals of every substep and every photon's delay from the same
keyed Philox streams with the same FP64 operation sequence, so labels, times, step rows and draw identities agree
exactly. The cases cover both-end and reflector sectors (reflector +1 and -1, whose SiPM s lands on channel s + 1),
speed and attenuation ratios, per-channel yield deviations, dead channels, steps with a NaN time, Cerenkov light
(charged particles above threshold), the Gauss branch of the photon numbers and 0.1 mm steps whose substeps draw
thousands of arrivals (the chunked Binomial). charges(tabs=...) then equals the oracle's SiPM stage on the oracle's
photons.
"""
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, lookup as LK, photons as PH  # noqa: E402
from oracle import lookup_optics as LO, response as RS  # noqa: E402

import lookup_fixtures as FX  # noqa: E402

SEED = 4242


def _setup(seed=0, n=700, nan=0.03, short=0.08, speed=None, atten=None):
    rng = np.random.default_rng(seed)
    G = FX.geometry()
    both = FX.random_table(rng, "LookupTable_both", reflector=0)
    refl = FX.random_table(rng, "LookupTable_refl", reflector=1)
    tabs = {"LookupTable_both": both, "LookupTable_refl": refl}
    sectors = {"EX": CF.SectorConfig("both", 0, 39794.0), "T1": CF.SectorConfig("refl", -1, 28610.0),
               "M1": CF.SectorConfig("refl", 1, 28610.0)}
    rc = CF.ResponseConfig(photon_yield_scale=0.94, photon_yield_variation_scale=1.0, sectors=sectors,
                           speed_ratio=speed or {"EX": 1.08}, atten_ratio=atten or {"T1": 0.6, "M1": 1.3})
    cfg = CF.ChainConfig(response=rc)
    cols, pdg = FX.steps(rng, G, n, n_events=9, short=short, nan=nan)
    cond = DT.nominal_conditions(4 * len(G.sector))
    cond = replace(cond, deviation=rng.normal(0, 0.05, len(cond.status)))
    st = np.array(cond.status)
    st[rng.choice(len(st), 6, replace=False)] |= 1
    st[rng.choice(len(st), 6, replace=False)] |= 4
    st[rng.choice(len(st), 6, replace=False)] |= 2      # bit 1 alone does not mask the SiPM stage
    cond = replace(cond, status=st)
    return G, tabs, cfg, cols, pdg, cond


def _oracle(G, tabs, cfg, cols, pdg, cond, seed, order, status=None):
    bars, tl = PH.bar_arrays(G, cfg.response, tabs)
    props = [LO.table_propagation(T) for T in tl]
    m, q, _ = PH.particle_props(pdg[order])
    c = {k: np.asarray(v)[order] for k, v in cols.items()}
    st = {"evkey": CH.event_key(c), "bar": c["bar"], "t0": c["t0"], "t1": c["t1"],
          "pos0": np.column_stack([c[f"{a}0"] for a in "xyz"]), "pos1": np.column_stack([c[f"{a}1"] for a in "xyz"]),
          "edep": c["edep"], "len": c["len"], "p0": np.sqrt(c["px0"] ** 2 + c["py0"] ** 2 + c["pz0"] ** 2),
          "p1": c["p1"], "mass": m, "charge": q}
    status = DT.effective_status(cond) if status is None else status
    return LO.photons(st, bars, tl, props, cond.deviation, status, cfg.response.photon_yield_scale,
                      cfg.response.photon_yield_variation_scale, seed)


def _device(G, tabs, cfg, cols, pdg, cond, seed):
    ctx = CH.context(cfg, cond)
    PH.upload(ctx, tabs, G, cfg.response)
    ls = PH.LookupSteps(CH.event_key(cols), cols, pdg)
    ctx.set_params(cfg, seed, ls.evkey_hi)
    pb, po = PH.device_photons(ctx, ls, cap_hint=16)   # too small: the library reports the size, the call is redone
    n = po.n
    return {k: CH._host(pb[k], n) for k in ("label", "time", "photon", "j", "rank")}, ls, po


def test_lookup_photons_bit_exact_vs_oracle():
    G, tabs, cfg, cols, pdg, cond = _setup()
    D, ls, po = _device(G, tabs, cfg, cols, pdg, cond, SEED)
    O = _oracle(G, tabs, cfg, cols, pdg, cond, SEED, ls.order)
    print(f"lookup photons: {len(O['time'])} oracle, {len(D['time'])} HIP; types {np.bincount(O['typ'])}")
    assert len(O["time"]) > 20000
    assert np.bincount(O["typ"], minlength=3).min() > 100            # scintillation and both Cerenkov kinds
    assert (O["chan"] % 4 == 3).any() and (O["chan"] % 4 == 1).any()  # reflector -1 writes SiPM 0 / 2 to 1 / 3
    assert np.array_equal(D["label"], O["evkey"] * CH.CHAN_MULT + O["chan"])
    assert np.array_equal(D["time"], O["time"])
    assert np.array_equal(D["photon"], O["row"])
    assert np.array_equal(D["j"].astype(np.int64), O["j"]) and np.array_equal(D["rank"].astype(np.int64), O["rank"])
    assert po.capped == O["capped"]
    dead = (DT.effective_status(cond)[O["chan"]] & 0b101) != 0
    assert not dead.any()


def test_lookup_photons_with_other_ratios_and_no_deviation():
    G, tabs, cfg, cols, pdg, cond = _setup(seed=5, n=300, nan=0.0, short=0.2, speed={"TAG": 0.9, "T1": 1.2},
                                           atten={"EX1": 0.7})
    cond = replace(cond, deviation=np.zeros(len(cond.status)))
    D, ls, _ = _device(G, tabs, cfg, cols, pdg, cond, 9)
    O = _oracle(G, tabs, cfg, cols, pdg, cond, 9, ls.order)
    assert np.array_equal(D["label"], O["evkey"] * CH.CHAN_MULT + O["chan"]) and np.array_equal(D["time"], O["time"])
    assert np.array_equal(D["photon"], O["row"]) and np.array_equal(D["rank"].astype(np.int64), O["rank"])


def test_make_photons_returns_every_channel_like_the_reference():
    G, tabs, cfg, cols, pdg, cond = _setup(seed=2, n=200)
    s, c, t = PH.make_photons(cols, pdg, G, cfg, tabs, cond=cond, key=SEED, evkey=CH.event_key(cols))
    order = np.argsort(CH.event_key(cols), kind="stable")
    O = _oracle(G, tabs, cfg, cols, pdg, cond, SEED, order, status=np.zeros(len(cond.status), np.int64))
    assert np.array_equal(s, order[O["row"]]) and np.array_equal(c, O["chan"]) and np.array_equal(t, O["time"])


def test_charges_with_tables_equal_the_oracle_sipm_stage():
    G, tabs, cfg, cols, pdg, cond = _setup(seed=3, n=500)
    # (thermal noise and trap afterpulses are refused by charges(), as by the reference's response/contract.py)
    cfg = replace(cfg, sipm=replace(cfg.sipm, cross_talk=0.1), window=(0.0, 1400.0))
    Q = CH.charges(cols, None, G, cond, cfg, None, key=SEED, tabs=tabs, pdg=pdg)
    order = np.argsort(CH.event_key(cols), kind="stable")
    O = _oracle(G, tabs, cfg, cols, pdg, cond, SEED, order)
    label = O["evkey"] * CH.CHAN_MULT + O["chan"]
    src = RS.PhiloxSource(SEED, (O["evkey"], O["chan"], O["j"], O["rank"]))
    R = RS.sipm_simulate(label, O["time"], order[O["row"]], cfg.photon_map, cfg.sipm, src, t_end=cfg.window[1])
    print(f"charges: {len(Q['time'])} HIP, {len(R['time'])} oracle from {len(label)} photons")
    assert len(Q["time"]) > 5000
    for k in ("chan", "time", "charge", "pe", "photon"):
        assert np.array_equal(Q[k], R[k]), k


def test_charges_with_tables_from_parts_and_edge_cases():
    G, tabs, cfg, cols, pdg, cond = _setup(seed=4, n=50)
    parts = {"run": cols["run"], "subrun": cols["subrun"], "event": cols["event"], "key": cols["sp"], "pdg": pdg}
    A = CH.charges(cols, parts, G, cond, cfg, None, key=3, tabs=tabs)
    B = CH.charges(cols, None, G, cond, cfg, None, key=3, tabs=tabs, pdg=pdg)
    assert all(np.array_equal(A[k], B[k]) for k in A)
    empty = {k: v[:0] for k, v in cols.items()}
    E = CH.charges(empty, None, G, cond, cfg, None, key=3, tabs=tabs, pdg=pdg[:0])
    assert all(len(v) == 0 for v in E.values())
    nan = dict(cols, t1=np.full(len(cols["t1"]), np.nan))
    N = CH.charges(nan, None, G, cond, cfg, None, key=3, tabs=tabs, pdg=pdg)
    assert all(len(v) == 0 for v in N.values())
    with pytest.raises(ValueError, match="CRVGeometry"):
        CH.charges(cols, None, DT.sector_geom(), cond, cfg, None, key=3, tabs=tabs, pdg=pdg)
    with pytest.raises(ValueError, match="bar"):
        CH.charges(dict(cols, bar=cols["bar"] + 1000), None, G, cond, cfg, None, key=3, tabs=tabs, pdg=pdg)
    hot = dict(tabs)
    hot["LookupTable_both"] = replace(tabs["LookupTable_both"], fiber_decay_time=8.7)
    with pytest.raises(ValueError, match="decay"):
        CH.charges(cols, None, G, cond, cfg, None, key=3, tabs=hot, pdg=pdg)
