"""HIP kernels (through the C-ABI) against the oracle on the same seeded inputs: each stage, then the whole chain.

Integer results (PE counts, charge rows, digi ADC words, cluster membership) must be bit-identical. The FP64
kernels evaluate the oracle's operation sequence, so pulse heights and times are compared with the oracle at 1e-12
relative; against the reference's recorded fits, times agree within 1e-9 relative and PE and height within 1e-9 or
at an end point whose chi2 equals the reference's within double-precision rounding (tests/fitcheck.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, synth as SY, template as TP  # noqa: E402
from oracle import chain as OC, pulsefit as OPF  # noqa: E402

import fitcheck  # noqa: E402

RTOL = 1e-12


@pytest.fixture(scope="module")
def setup():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    tmpl = TP.load(cfg.waveform)
    S = SY.cosmic_events(300, seed=0)
    Qo = OC.charges(S, cond, cfg, 5)
    return cfg, cond, G, tmpl, S, Qo


def test_charges_bit_exact(setup):
    cfg, cond, G, tmpl, S, Qo = setup
    timers = {}
    Q = CH.charges(S, None, G, cond, cfg, None, key=5, timers=timers)
    assert timers["photons"] == Qo["n_photons"]
    for k in ("chan", "time", "photon", "pe", "charge"):
        np.testing.assert_array_equal(Q[k], Qo[k], err_msg=k)


def test_digis_and_records_bit_exact(setup):
    cfg, cond, G, tmpl, S, Qo = setup
    D = CH.digis(Qo, cond, cfg, None, key=5)
    Do = OC.digis(Qo, cond, cfg, 5, tmpl)
    assert len(D["chan"]) == len(Do["chan"]) > 1000
    for k in ("chan", "starttdc", "adc", "start", "tdc0"):
        np.testing.assert_array_equal(D[k], Do[k], err_msg=k)
    np.testing.assert_array_equal(D["sample"], Do["digi_sample"])
    np.testing.assert_array_equal(D["record_adc"], Do["record_adc"].astype(np.uint16))
    assert D["overlong"] == Do["overlong"]


def test_reco_waveforms_match(setup):
    cfg, cond, G, tmpl, S, Qo = setup
    Do = OC.digis(Qo, cond, cfg, 5, tmpl)
    W = CH.reco_waveforms(Do, cond)
    Wo = OC.reco_waveforms(Do, cond)
    for k in ("ev", "bar", "sipm", "chan", "n", "start", "adc", "ped", "area", "height", "toff"):
        np.testing.assert_array_equal(W[k], Wo[k], err_msg=k)


def _pulse_compare(P, Po):
    for k in ("wf", "seq", "failed", "zerondf", "pedfromdb"):
        np.testing.assert_array_equal(P[k], Po[k], err_msg=k)
    for k in ("pe", "peph", "time", "height", "beta", "chi2", "letime", "ped", "fit_A", "fit_mu", "fit_b"):
        np.testing.assert_allclose(P[k], Po[k], rtol=RTOL, atol=0, err_msg=k)


def test_pulses_match(setup):
    cfg, cond, G, tmpl, S, Qo = setup
    Wo = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 5, tmpl), cond)
    P = CH.pulses(Wo, cond, cfg, G)
    _pulse_compare(P, OC.pulses(Wo, cond, cfg))


def test_pulses_match_reference_golden(golden):
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    for name in ("synth12", "edge36", "double36", "train252", "chain"):
        args = [golden[f"pf_{name}_in_{k}"] for k in ("adc", "n", "start", "ped", "area", "height")]
        n = len(args[1])
        W = dict(ev=np.zeros(n, np.int64), bar=np.zeros(n, np.int64), sipm=np.zeros(n, np.int64), adc=args[0],
                 n=args[1], start=args[2], ped=args[3], area=args[4], height=args[5], toff=np.zeros(n))
        P = CH.pulses(W, cond, cfg)
        win = []
        _pulse_compare(P, OPF.reconstruct_full(*args, windows=win))
        g = lambda k: golden[f"pf_{name}_{k}"]
        for k in ("wf", "seq", "failed", "zerondf", "pedfromdb"):
            np.testing.assert_array_equal(P[k], g(k), err_msg=k)
        # times within 1e-9 relative; PE and height within 1e-9 or chi2-equivalent to the reference's end point
        fitcheck.check_fit_values(P, {k: g(k) for k in fitcheck.BOUNDS.keys() | {"wf", "seq", "time", "letime",
                                                                                   "fit_mu"}}, win)


def _cluster_compare(out, ref):
    for k in ("cl_event", "cl_type", "cl_poscalc", "mem_cl", "mem_idx", "mem_event"):
        np.testing.assert_array_equal(np.asarray(out[k]).astype(np.int64), np.asarray(ref[k]).astype(np.int64),
                                      err_msg=k)
    for k in ("cl_pe", "cl_start", "cl_end", "cl_avgtime", "cl_pos"):
        np.testing.assert_allclose(out[k], ref[k], rtol=RTOL, atol=1e-9, err_msg=k)
    np.testing.assert_allclose(out["cl_slope"], ref["cl_slope"], rtol=1e-9, atol=1e-9)


def test_clusters_match(setup):
    cfg, cond, G, tmpl, S, Qo = setup
    Wo = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 5, tmpl), cond)
    out, _ = CH.clusters(Wo, cond, G, None, cfg, detail=True)
    ref, _ = OC.clusters(Wo, cond, G, cfg)
    assert len(ref["cl_event"]) > 100
    _cluster_compare(out[0], ref)


def test_clusters_match_in_every_event_size_class(setup):
    """Events of up to 128 and up to 256 pulses take the wavefront finders, larger ones the block finder: merge
    groups of eight events (about 380 pulses) and of four (about 190), keep the rest single; all must equal the
    oracle."""
    cfg, cond, G, tmpl, S, Qo = setup
    Wo = dict(OC.reco_waveforms(OC.digis(Qo, cond, cfg, 5, tmpl), cond))
    ev = np.asarray(Wo["ev"])
    Wo["ev"] = np.where(ev < 80, ev // 8, np.where(ev < 160, 100 + ev // 4, ev + 1000))
    out, _ = CH.clusters(Wo, cond, G, None, cfg, detail=True)
    ref, _ = OC.clusters(Wo, cond, G, cfg)
    npl = np.bincount(np.unique(ref["pulses"]["ev"], return_inverse=True)[1])
    assert (npl > 256).any() and ((npl > 128) & (npl <= 256)).any() and (npl <= 128).any()
    _cluster_compare(out[0], ref)


def _dense_pulses(rng, n_events):
    """Pulse tables of shower-like events: bars of 3 to 24 adjacent columns over two modules fire on every layer
    within a few ns, four SiPMs each, with secondary and late pulses (late ones test the time window hit by hit).
    Rows sorted by (event, bar, SiPM, time) with seq the pulse's rank on its SiPM."""
    rows = []
    for e in range(n_events):
        ncol = int(rng.integers(3, 25))
        c0 = int(rng.integers(0, 2 * DT.N_COLS - ncol + 1))
        m0 = int(rng.integers(0, DT.N_MODULES - 1))
        occ = rng.uniform(0.4, 1.0)
        t0 = rng.uniform(100.0, 1500.0)
        for layer in range(DT.N_LAYERS):
            for gc in range(c0, c0 + ncol):
                if rng.random() > occ:
                    continue
                bar = int(DT.counter_id(m0 + gc // DT.N_COLS, layer, gc % DT.N_COLS))
                for sp in range(4):
                    if rng.random() < 0.1:
                        continue
                    ts = [t0 + rng.normal(0.0, 2.0)]
                    if rng.random() < 0.15:
                        ts.append(ts[0] + rng.uniform(8.0, 20.0))
                    if rng.random() < 0.1:
                        ts[0] += rng.uniform(6.0, 14.0)
                    for k, tt in enumerate(sorted(ts)):
                        rows.append((e, bar, sp, tt, rng.uniform(3.0, 40.0), k))
    a = np.array(rows)
    o = np.lexsort((a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    a = a[o]
    return (a[:, 0].astype(np.int64), a[:, 1].astype(np.int64), a[:, 2].astype(np.int64), a[:, 3].copy(),
            a[:, 4].copy(), a[:, 5].astype(np.int64))


def test_clusters_match_on_dense_events():
    """Large clusters (up to about 10^5 layer combinations on a side) in events of every finder size class."""
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    ev, bar, sipm, time, pe, seq = _dense_pulses(np.random.default_rng(21), 24)
    npl = np.bincount(ev)
    assert (npl > 256).any() and ((npl > 128) & (npl <= 256)).any() and (npl <= 128).any()
    out = CH.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    ref = OC.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    assert len(ref["cl_event"]) >= 20
    _cluster_compare(out, ref)


def test_clusters_match_on_merged_dense_events():
    """Two events of twenty showers each (1,400-1,900 hits on a side): lists longer than the block finder's LDS stage,
    clustered in cell order with a union-find over the links, dense components included; all must equal the
    oracle."""
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    ev, bar, sipm, time, pe, _ = _dense_pulses(np.random.default_rng(33), 40)
    ev = ev // 20
    o = np.lexsort((time, sipm, bar, ev))
    ev, bar, sipm, time, pe = ev[o], bar[o], sipm[o], time[o], pe[o]
    key = (ev * DT.N_COUNTERS + bar) * 4 + sipm
    start = np.r_[0, np.flatnonzero(np.diff(key)) + 1]
    seq = np.arange(len(key)) - np.repeat(start, np.diff(np.r_[start, len(key)]))  # rank on its SiPM
    for side in (0, 1):
        assert (np.bincount(ev[sipm % 2 == side]) > 1300).all()
    out = CH.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    ref = OC.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    assert len(ref["cl_event"]) >= 30
    _cluster_compare(out, ref)


def test_clusters_match_on_eighty_merged_showers():
    """One event of eighty showers (6,400 hits on a side): longer than the block finder's union-find fits beside the
    cell bins in LDS, so its labels propagate in rounds over the cell-order lists; must equal the oracle (the oracle
    takes about a minute here)."""
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    ev, bar, sipm, time, pe, _ = _dense_pulses(np.random.default_rng(35), 80)
    ev = np.zeros_like(ev)
    o = np.lexsort((time, sipm, bar))
    bar, sipm, time, pe = bar[o], sipm[o], time[o], pe[o]
    key = bar * 4 + sipm
    start = np.r_[0, np.flatnonzero(np.diff(key)) + 1]
    seq = np.arange(len(key)) - np.repeat(start, np.diff(np.r_[start, len(key)]))
    for side in (0, 1):
        assert (sipm % 2 == side).sum() > 6000
    out = CH.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    ref = OC.find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    assert len(ref["cl_event"]) >= 50
    _cluster_compare(out, ref)


def test_clusters_match_reference_golden(golden):
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    for name in ("synth", "chain"):
        g = lambda k: golden[f"co_{name}_{k}"]
        status = g("status").astype(np.int64)
        cond_s = DT.Conditions(cond.pedestal, cond.pulsearea, cond.pulseheight, cond.timeoffset, status, cond.roc,
                               cond.feb, cond.deviation)
        ev = g("in_ev")
        o = np.lexsort((g("in_idx"), ev))
        assert np.array_equal(o, np.arange(len(o)))
        out = CH.find_clusters(ev, g("in_bar"), g("in_sipm"), g("in_time"), g("in_pe"), g("in_seq"), cfg, cond_s, G)
        np.testing.assert_array_equal(out["cl_event"], g("cl_event"))
        np.testing.assert_array_equal(out["mem_idx"], g("members"))
        np.testing.assert_allclose(out["cl_pe"], g("cl_pe"), rtol=1e-12)


def test_chain_edge_cases():
    cfg = CF.ChainConfig(sipm=CF.SiPMConfig(cross_talk=0.2, pixel_id="truncate"),
                         photon_map=np.random.default_rng(3).random((40, 40)) ** 2)
    cond = DT.nominal_conditions()
    cond.status[[0, 9, 130]] = [1, 4, 2]          # bit 0 / 2: no SiPM; bit 1: pulses ignored
    G = DT.sector_geom()
    tmpl = TP.load(cfg.waveform)
    S = SY.uneven_events(60, seed=4)
    Q = CH.charges(S, None, G, cond, cfg, None, key=9)
    Qo = OC.charges(S, cond, cfg, 9)
    for k in ("chan", "time", "photon", "pe"):
        np.testing.assert_array_equal(Q[k], Qo[k], err_msg=k)
    D = CH.digis(Qo, cond, cfg, None, key=9)
    Do = OC.digis(Qo, cond, cfg, 9, tmpl)
    np.testing.assert_array_equal(D["adc"], Do["adc"])
    Wo = OC.reco_waveforms(Do, cond)
    _pulse_compare(CH.pulses(Wo, cond, cfg, G), OC.pulses(Wo, cond, cfg))


def test_overlay_events_stage_by_stage():
    """Two events of the overlay workload (cosmic mixture + 0.5 background hits per counter, ~2,800 steps each):
    charges and digis bit-exact, pulses and clusters against the oracle on the same reco waveforms."""
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    tmpl = TP.load(cfg.waveform)
    S = SY.overlay_events(2, seed=6)
    Q = CH.charges(S, None, G, cond, cfg, None, key=13)
    Qo = OC.charges(S, cond, cfg, 13)
    assert len(Qo["chan"]) > 100_000
    for k in ("chan", "time", "photon", "pe"):
        np.testing.assert_array_equal(Q[k], Qo[k], err_msg=k)
    D = CH.digis(Qo, cond, cfg, None, key=13)
    Do = OC.digis(Qo, cond, cfg, 13, tmpl)
    np.testing.assert_array_equal(D["adc"], Do["adc"])
    Wo = OC.reco_waveforms(Do, cond)
    _pulse_compare(CH.pulses(Wo, cond, cfg, G), OC.pulses(Wo, cond, cfg))
    out, _ = CH.clusters(Wo, cond, G, None, cfg, detail=True)
    ref, _ = OC.clusters(Wo, cond, G, cfg)
    _cluster_compare(out[0], ref)


def test_empty_inputs():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    G = DT.sector_geom()
    S = SY.cosmic_events(0, seed=0)
    Q = CH.charges(S, None, G, cond, cfg, None, key=1)
    assert len(Q["chan"]) == 0
    D = CH.digis(Q, cond, cfg, None, key=1)
    assert len(D["chan"]) == 0
    W = CH.reco_waveforms(D, cond)
    out, timers = CH.clusters(W, cond, G, None, cfg, detail=True)
    assert len(out[0]["cl_event"]) == 0 and timers["pulses"] == 0


def test_clusters_with_background_pulses_match(setup):
    """cfg.background: background pulses drawn for every event join the signal pulses in (event, bar, SiPM, time)
    order before the finder (reference chain.py:337-342); the HIP finder on the merged table equals the oracle's."""
    from dataclasses import replace
    from crvhip import background as BG
    cfg, cond, G, tmpl, S, Qo = setup
    rng = np.random.default_rng(17)
    sizes = rng.integers(1, 4, 400)
    grp = np.repeat(np.arange(400), sizes)
    n = len(grp)
    bars = np.flatnonzero(np.asarray(G.sector) >= 0)
    lib = {"group": grp, "bar": rng.choice(bars, n), "sipm": rng.integers(0, 4, n), "pe": rng.uniform(2, 40, n),
           "dt": np.concatenate([np.r_[0.0, np.sort(rng.uniform(0, 25, k - 1))] for k in sizes]),
           "seq": np.zeros(n, np.int64)}
    M = BG.BackgroundModel.from_partials([dict(events=1000, exposure_ns=1000 * 1e5, groups=3000, cnt_sum=3000,
                                               cnt_sumsq=3000 + 9000 + 4500, lib=lib)])
    cfgb = replace(cfg, background=M, background_scale=1.0)
    Wo = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 5, tmpl), cond)
    timers = {}
    out, _ = CH.clusters(Wo, cond, G, None, cfgb, timers=timers, detail=True, rng=np.random.default_rng(9))
    assert timers["background_pulses"] > 100
    P = out[0]["pulses"]
    ref = OC.find_clusters(P["ev"], P["bar"], P["sipm"], P["time"], P["pe"], P["seq"], cfg, cond, G)
    _cluster_compare(out[0], ref)
    # the same draws as the sampler on its own: the merged table holds the signal pulses plus exactly these
    sig, _ = CH.clusters(Wo, cond, G, None, cfg, detail=True)
    B = M.sample(np.unique(np.asarray(Wo["ev"])), cfg.window, np.random.default_rng(9))
    assert len(P["ev"]) == len(sig[0]["pulses"]["ev"]) + len(B["ev"])
