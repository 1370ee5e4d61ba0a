"""The staged onspill response (crvhip.onspill: crv_photons_run with the microbunch fold, then crv_sipm_run with dark
noise on every live channel) against the oracle's restatement of the reference's onspill.py, through the C-ABI.

Both sides draw the photons and the dark noise from the same keyed Philox streams and share the FP64 arithmetic, so
the charge tables agree bit for bit. Without dark noise and bunch shifts, the staged path equals the fused charges
stage (crv_charges_run) on the same steps, which pins the photon list's order and draw identity.
"""
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, onspill as O, readout as RO, synth as SY  # noqa: E402
from oracle import onspill as OO  # noqa: E402

SEED = 77


def _population(n_events, n_steps_events, seed, spread=2.0):
    """cosmic steps of the first n_steps_events events (t0 / t1 around the step time), keys (1, 0, e) for every
    event, random bunch times."""
    S = SY.cosmic_events(n_steps_events, seed=seed)
    rng = np.random.default_rng(seed)
    n = len(S["event"])
    steps = {"run": np.ones(n, np.int64), "subrun": np.zeros(n, np.int64), "event": S["event"].astype(np.int64),
             "counter": S["counter"], "edep": S["edep"], "z": S["z"],
             "t0": S["t"] - spread, "t1": S["t"] + spread}
    keys = np.column_stack([np.ones(n_events, np.int64), np.zeros(n_events, np.int64), np.arange(n_events)])
    tm = O.EventTiming(keys=keys, mc_pbtime=rng.uniform(-2500, 2500, n_events),
                       reco_pbtime=rng.uniform(-50, 50, n_events), event_length=np.full(n_events, 1675.))
    return steps, tm


def _cond(masked_frac=0.05, seed=3):
    cond = DT.nominal_conditions()
    rng = np.random.default_rng(seed)
    mask = rng.random(len(cond.status)) >= masked_frac
    st = np.array(cond.status)
    st[rng.choice(len(st), 200, replace=False)] = rng.choice([1, 2, 4], 200)
    return replace(cond, status=st, readout_mask=mask)


def _oracle(steps, tm, G, cond, cfg):
    ev = tm.lookup(np.column_stack([steps[k] for k in ("run", "subrun", "event")]))
    status = RO.effective_status(cond)
    dev = (np.asarray(cond.deviation) * cfg.response.photon_yield_variation_scale + 1.0) * cfg.response.photon_yield_scale
    t = 0.5 * (steps["t0"] + steps["t1"])
    return OO.charges(ev, steps["counter"], steps["edep"], steps["z"], t, steps["t0"], steps["t1"], status, dev,
                      cfg.optics, cfg.sipm, cfg.photon_map, SEED, tm.mc_pbtime, tm.period, cfg.window,
                      len(tm.keys), len(G.sector) * 4)


def _same(Q, R):
    assert len(Q["chan"]) == len(R["chan"])
    for k in ("chan", "photon", "time", "pe", "charge"):
        np.testing.assert_array_equal(Q[k], R[k], err_msg=k)


def test_onspill_charges_match_oracle():
    """150 events, 100 of them with cosmic steps, bunch times of up to +-2500 ns (steps and photons fold across the
    period), 5 % of the sensors absent, dead and reco-only channels, dark noise at 2e-5 per ns."""
    steps, tm = _population(150, 100, seed=8)
    G, cond = DT.sector_geom(), _cond()
    cfg = replace(CF.ChainConfig(), window=(200, 1500), sipm=replace(CF.SiPMConfig(), thermal_rate=2e-5))
    timers = {}
    Q = O.charges(steps, np.full(len(steps["event"]), 13), G, cond, cfg, None, timing=tm, key=SEED, timers=timers)
    R = _oracle(steps, tm, G, cond, cfg)
    _same(Q, R)
    dark = Q["photon"] == -1
    assert dark.sum() > 1000 and (~dark).sum() > 1000
    assert timers["time_gate_steps_skipped"] > 0
    assert np.all(Q["time"][dark] >= 100) and np.all(Q["time"][dark] <= 1500)
    assert np.all(Q["time"] >= 0) and np.all(Q["time"] <= 1500)
    live = (RO.effective_status(cond)[Q["chan"] % CH.CHAN_MULT] & 0b101) == 0
    assert live.all()
    assert np.isin(np.arange(100, 150), Q["chan"] // CH.CHAN_MULT).all()  # events without steps: dark noise


def test_staged_path_equals_fused_charges_without_noise_and_shifts():
    """mc_pbtime 0 and no dark noise: photons list -> SiPM stage gives the fused stage's charges (same labels when
    the event keys are the dense rows)."""
    S = SY.cosmic_events(300, seed=5)
    n = len(S["event"])
    t = 200.0 + np.fmod(S["t"], 1000.0)
    steps = {"run": np.zeros(n, np.int64), "subrun": np.zeros(n, np.int64), "event": S["event"].astype(np.int64),
             "counter": S["counter"], "edep": S["edep"], "z": S["z"], "t0": t, "t1": t}
    keys = np.column_stack([np.zeros(300, np.int64), np.zeros(300, np.int64), np.arange(300)])
    tm = O.EventTiming(keys=keys, mc_pbtime=np.zeros(300), reco_pbtime=np.zeros(300), event_length=np.full(300, 1675.))
    G, cond = DT.sector_geom(), _cond()
    cfg = replace(CF.ChainConfig(), window=(200, 1500))
    Q = O.charges(steps, np.full(n, 13), G, cond, cfg, None, timing=tm, key=SEED)
    F = CH.charges({"event": S["event"], "counter": S["counter"], "edep": S["edep"], "z": S["z"], "t": t}, None, G,
                   cond, cfg, None, key=SEED)
    _same(Q, F)


def test_zero_step_events_get_dark_noise_on_live_channels_only():
    tm = O.EventTiming(keys=np.column_stack([np.ones(32, int), np.zeros(32, int), np.arange(32)]),
                       mc_pbtime=np.zeros(32), reco_pbtime=np.zeros(32), event_length=np.full(32, 1675.))
    G, cond = DT.sector_geom(), _cond(0.3)
    cfg = replace(CF.ChainConfig(), window=(200, 1500), sipm=replace(CF.SiPMConfig(), thermal_rate=1e-4))
    empty = {k: np.zeros(0, np.int64 if k in ("run", "subrun", "event", "counter") else float)
             for k in ("run", "subrun", "event", "counter", "edep", "z", "t0", "t1")}
    Q = O.charges(empty, np.zeros(0, int), G, cond, cfg, None, timing=tm, key=SEED)
    assert len(Q["time"]) > 10000
    assert np.all(Q["photon"] == -1)
    assert np.all((Q["time"] >= 100) & (Q["time"] <= 1500))
    assert np.all((RO.effective_status(cond)[Q["chan"] % CH.CHAN_MULT] & 0b101) == 0)
    np.testing.assert_array_equal(np.lexsort((Q["time"], Q["chan"])), np.arange(len(Q["time"])))
    _same(Q, _oracle(empty, tm, G, cond, cfg))


def test_run_applies_the_reconstructed_bunch_correction():
    """run(): digis, waveforms and clusters downstream of the onspill charges; every waveform carries -reco_pbtime of
    its event and moving reco_pbtime moves every pulse and cluster time by exactly that much."""
    steps, tm = _population(60, 60, seed=12)
    G, cond = DT.sector_geom(), _cond(0.0)
    cfg = replace(CF.ChainConfig(), window=(200, 1500), sipm=replace(CF.SiPMConfig(), thermal_rate=1e-5))
    pdg = np.full(len(steps["event"]), 13)
    a = O.run(steps, pdg, G, cond, cfg, None, timing=tm, counters_cfg=None, key=SEED, detail=True)
    tm2 = replace(tm, reco_pbtime=tm.reco_pbtime + 64.0)
    b = O.run(steps, pdg, G, cond, cfg, None, timing=tm2, counters_cfg=None, key=SEED, detail=True)
    np.testing.assert_array_equal(a["W"]["event_t0"], -tm.reco_pbtime[a["W"]["ev"]])
    np.testing.assert_array_equal(a["event_keys"], tm.keys)
    _same(a["Q"], b["Q"])
    ca, cb = a["clusters"][0], b["clusters"][0]
    assert len(ca["cl_event"]) > 20
    for k in ("cl_event", "mem_idx", "mem_cl"):
        np.testing.assert_array_equal(ca[k], cb[k], err_msg=k)
    np.testing.assert_allclose(cb["pulses"]["time"], ca["pulses"]["time"] - 64.0, rtol=0, atol=1e-9)
    np.testing.assert_allclose(cb["cl_start"], ca["cl_start"] - 64.0, rtol=0, atol=1e-9)
