"""The reference's stage contracts re-expressed against crvhip.chain on the device (crv/tests/test_chain.py:18-35,
crv/tests/test_chain_reco_contract.py:51-56), and the reference-geometry and photon-map paths of the boundary."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, synth as SY, template as TP  # noqa: E402
from oracle import chain as OC  # noqa: E402


def _cond(n=16):
    """Reference crv/tests/test_chain.py:13-16: 4 channels per FEB."""
    z = np.zeros(n)
    feb = np.repeat(np.arange(n // 4), 4)
    return DT.Conditions(z + 100, z + 400, z + 11, z, np.zeros(n, np.int64), np.ones(n, np.int64), feb, z)


def test_tdc0_one_draw_per_event_and_feb():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    Q = CH.charges(SY.cosmic_events(40, seed=3), None, None, cond, cfg, None, key=2)
    D = CH.digis(Q, cond, cfg, None, key=2)
    ev, ch = D["chan"] // CF.CHAN_MULT, D["chan"] % CF.CHAN_MULT
    k = ev * 1_000_000 + cond.roc[ch] * 1000 + cond.feb[ch]
    u, inv = np.unique(k, return_inverse=True)
    first = np.zeros(len(u))
    first[inv] = D["tdc0"]
    np.testing.assert_array_equal(D["tdc0"], first[inv])           # one value per (event, ROC, FEB)
    assert len(np.unique(first)) == len(u) > 20                     # distinct draws
    assert np.all((D["tdc0"] >= 0) & (D["tdc0"] < CF.PERIOD))


def test_reco_waveforms_concatenates_consecutive_digis():
    cond = _cond()
    CH.context(CF.ChainConfig(), cond)
    lab = np.array([5 * CF.CHAN_MULT + 2] * 3)
    D = {"chan": lab, "starttdc": np.array([10, 22, 40]), "adc": np.arange(36).reshape(3, 12) + 100}
    W = CH.reco_waveforms(D, cond)
    assert W["n"].tolist() == [24, 12] and W["start"].tolist() == [10, 40] and W["ev"].tolist() == [5, 5]
    assert W["bar"].tolist() == [0, 0] and W["sipm"].tolist() == [2, 2] and np.all(W["ped"] == 100)
    np.testing.assert_array_equal(W["adc"][0, :24], np.arange(24) + 100)


def test_reco_waveforms_carries_offsets():
    z = np.zeros(8)
    cond = DT.Conditions(z + 100, z + 400, z + 11, np.arange(8) * 7.0 - 13, np.arange(8, dtype=np.int64),
                         z.astype(np.int64), z.astype(np.int64), z)
    CH.context(CF.ChainConfig(), cond)
    ch = np.array([3, 0, 5])
    D = {"chan": CF.CHAN_MULT + ch, "starttdc": np.zeros(3, np.int64), "adc": np.full((3, 12), 100)}
    W = CH.reco_waveforms(D, cond)
    np.testing.assert_array_equal(W["toff"], cond.timeoffset[W["chan"]])


def test_clusters_with_a_crv_geometry_equal_the_synthetic_detector():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    tmpl = TP.load(cfg.waveform)
    Qo = OC.charges(SY.cosmic_events(120, seed=8), cond, cfg, 4)
    Wo = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 4, tmpl), cond)
    geom, counters = DT.synthetic_crv_geometry()
    a, _ = CH.clusters(Wo, cond, geom, counters, cfg, detail=True)
    b, _ = CH.clusters(Wo, cond, None, None, cfg, detail=True)
    assert len(b[0]["cl_event"]) > 50
    for k in b[0]:
        if k != "pulses":
            np.testing.assert_array_equal(a[0][k], b[0][k], err_msg=k)


def test_charges_from_world_coordinates_and_a_photon_map():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    S = SY.cosmic_events(30, seed=5)
    geom, _ = DT.synthetic_crv_geometry()
    bar = np.asarray(S["counter"])
    W = dict(subrun=np.zeros(len(bar), np.int64), event=np.asarray(S["event"]), bar=bar,
             edep=np.asarray(S["edep"]), t0=np.asarray(S["t"]), t1=np.asarray(S["t"]))
    for a, j in (("x", 0), ("y", 1), ("z", 2)):
        W[f"{a}0"] = W[f"{a}1"] = geom.center[bar, j] + (np.asarray(S["z"]) if j == 2 else 0.0)
    Q1 = CH.charges(W, None, geom, cond, cfg, None, key=3)
    Q0 = CH.charges(S, None, None, cond, cfg, None, key=3)
    for k in ("chan", "time", "photon", "pe", "charge"):
        np.testing.assert_array_equal(Q1[k], Q0[k], err_msg=k)
    pmap = np.random.default_rng(1).random((40, 40)) ** 3
    Qp = CH.charges(S, None, None, cond, cfg, None, key=3, pmap=pmap)
    from dataclasses import replace
    Qo = OC.charges(S, cond, replace(cfg, photon_map=pmap), 3)
    for k in ("chan", "time", "photon", "pe", "charge"):
        np.testing.assert_array_equal(Qp[k], Qo[k], err_msg=k)
    assert len(Qp["chan"]) != len(Q0["chan"]) or not np.array_equal(Qp["pe"], Q0["pe"])
    Q2 = CH.charges(S, None, None, cond, cfg, None, key=3)         # the configured map is back
    np.testing.assert_array_equal(Q2["pe"], Q0["pe"])


def _two_sector_geom():
    """The synthetic detector cut into two sectors of two sector types: modules 0-42 "A" (type 1, 3 of 4 layers) and
    43-85 "B" (type 2, 4 of 4 layers, 8 ns), B's counters numbered from 0 within its type."""
    G = DT.sector_geom()
    module = np.arange(DT.N_COUNTERS) // (DT.N_LAYERS * DT.N_COLS)
    col = np.arange(DT.N_COUNTERS) % DT.N_COLS
    sec = (module >= 43).astype(np.int64)
    cit = np.where(sec == 0, module * DT.N_COLS + col, (module - 43) * DT.N_COLS + col)
    G2 = DT.SectorGeom(sec, G.layer, cit.astype(np.int64), G.x, G.y, G.pos, np.array([2, 2]), np.array([1, 2]),
                       ["A", "B"], np.r_[G.pos_side0, G.pos_side0], np.r_[G.pos_side1, G.pos_side1],
                       np.array([True, True]), np.array([True, True]))
    from dataclasses import replace
    cuts = {"A": CF.SYNTH_CUTS, "B": replace(CF.SYNTH_CUTS, coincidence_layers=4, max_dt=8.0)}
    return G2, replace(CF.ChainConfig(), finder=replace(CF.FinderConfig(), cuts=cuts))


def test_clusters_match_on_two_sector_types():
    G2, cfg = _two_sector_geom()
    cond = DT.nominal_conditions()
    tmpl = TP.load(cfg.waveform)
    Qo = OC.charges(SY.uneven_events(150, seed=12), cond, cfg, 6)
    Wo = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 6, tmpl), cond)
    out, _ = CH.clusters(Wo, cond, G2, None, cfg, detail=True)
    ref, _ = OC.clusters(Wo, cond, G2, cfg)
    assert len(ref["cl_event"]) > 30 and len(np.unique(ref["cl_type"])) == 2
    for k in ("cl_event", "cl_type", "cl_poscalc", "mem_cl", "mem_idx", "mem_event"):
        np.testing.assert_array_equal(np.asarray(out[0][k]).astype(np.int64), np.asarray(ref[k]).astype(np.int64),
                                      err_msg=k)
    for k in ("cl_pe", "cl_start", "cl_end", "cl_avgtime", "cl_pos", "cl_slope"):
        np.testing.assert_allclose(out[0][k], ref[k], rtol=1e-12, atol=1e-9, err_msg=k)
