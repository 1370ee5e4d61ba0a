"""The stage functions fail closed, as the reference's do (crv/tests/test_response_contract.py:115-128,
crv/tests/test_chain.py:38-41): unsupported settings raise ValueError naming the setting before any device work, and
a reference CRVGeometry is either honoured or refused. Runs without a GPU."""
from dataclasses import replace

import numpy as np
import pytest

from crvhip import chain as CH, config as CF, contract as CT, detector as DT, synth as SY


@pytest.mark.parametrize("setting", ["thermal_rate", "trap0_prob", "trap1_prob"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_unsupported_processes_refused_before_response(setting, precision):
    cfg = CF.ChainConfig(precision=precision)
    cfg = replace(cfg, sipm=replace(cfg.sipm, **{setting: 0.01}))
    S = SY.cosmic_events(2, seed=0)
    cond = DT.nominal_conditions()
    with pytest.raises(ValueError, match=setting):
        CH.charges(S, None, None, cond, cfg, None, key=1)
    with pytest.raises(ValueError, match=setting):
        CH.run_chain(S, cond, cfg, key=1)
    with pytest.raises(ValueError, match=setting):
        CH.clusters({}, cond, None, None, cfg)


@pytest.mark.parametrize("window", [(0.0, np.nan), (100.0, 100.0), (200.0, 100.0), (0.0, np.inf), (0.0,)])
def test_invalid_window_refused(window):
    with pytest.raises(ValueError, match="window"):
        CT.validate_window(window)
    cfg = replace(CF.ChainConfig(), window=window)
    with pytest.raises(ValueError, match="window"):
        CH.charges(SY.cosmic_events(1, seed=0), None, None, DT.nominal_conditions(), cfg, None, key=1)


def test_pixel_id_refused():
    cfg = CF.ChainConfig()
    cfg = replace(cfg, sipm=replace(cfg.sipm, pixel_id="floor"))
    with pytest.raises(ValueError, match="pixel_id"):
        CH.charges(SY.cosmic_events(1, seed=0), None, None, DT.nominal_conditions(), cfg, None, key=1)


@pytest.mark.parametrize("arg", ["tabs", "pdg", "caps"])
def test_unhonoured_response_arguments_refused(arg):
    S = SY.cosmic_events(1, seed=0)
    with pytest.raises(ValueError, match=arg):
        CH.charges(S, None, None, DT.nominal_conditions(), CF.ChainConfig(), None, key=1, **{arg: object()})


def test_clusters_refuses_unknown_geometry_and_caps():
    cond, cfg = DT.nominal_conditions(), CF.ChainConfig()
    with pytest.raises(ValueError, match="geom"):
        CH.clusters({}, cond, object(), None, cfg)
    geom, _ = DT.synthetic_crv_geometry()
    with pytest.raises(ValueError, match="counters_cfg"):
        CH.clusters({}, cond, geom, None, cfg)
    with pytest.raises(ValueError, match="caps"):
        CH.clusters({}, cond, None, None, cfg, caps=object())


def test_response_refuses_counters_of_another_length():
    geom, _ = DT.synthetic_crv_geometry()
    geom.halflen[7, 2] = 1000.0
    S = SY.cosmic_events(1, seed=0)
    with pytest.raises(ValueError, match="half length"):
        CH.charges(S, None, geom, DT.nominal_conditions(), CF.ChainConfig(), None, key=1)


def test_sector_geom_from_crv_geometry_is_the_synthetic_finder_geometry():
    geom, counters = DT.synthetic_crv_geometry()
    G = DT.sector_geom_from(geom, counters)
    R = DT.sector_geom()
    for f in ("sector", "layer", "counter_in_type", "x", "y", "pos", "length_dir", "sector_type", "pos_side0",
              "pos_side1", "sipms_side0", "sipms_side1"):
        np.testing.assert_array_equal(getattr(G, f), getattr(R, f), err_msg=f)
    assert G.sector_names == R.sector_names


def test_sector_geom_from_counts_preceding_sectors():
    """Two sectors, B preceded by A: B's counters continue A's numbering within their sector type (reference
    reco/coinc.py:88-96), x / y follow each sector's own width / thickness axes, sides from the counters file."""
    nA, nB = 2 * 3, 1 * 2            # A: 2 modules x 3 counters; B: 1 module x 2 counters (one layer each)
    sector = np.r_[np.zeros(nA, int), np.ones(nB, int)]
    module = np.r_[np.repeat([0, 1], 3), [0, 0]]
    counter = np.r_[np.tile([0, 1, 2], 2), [0, 1]]
    center = np.array([[10.0 * i, 100.0 + i, 1000.0 - i] for i in range(nA + nB)])
    halflen = np.tile([5.0, 6.0, 700.0], (nA + nB, 1))
    l2w = np.r_[np.tile([1, 0, 2], (nA, 1)), np.tile([2, 1, 0], (nB, 1))]
    geom = DT.CRVGeometry(["A", "B"], sector, module, np.zeros(nA + nB, int), counter, center, halflen, l2w)
    cfgs = {"crs.sectorTypeA": 3, "crs.precedingSectorForA": -1, "crs.nCountersPerModuleA": 3, "crs.nModulesA": 2,
            "crs.sectorTypeB": 3, "crs.precedingSectorForB": 0, "crs.nCountersPerModuleB": 2, "crs.nModulesB": 1,
            "crs.sipmsAtSide1B": False}
    G = DT.sector_geom_from(geom, cfgs)
    np.testing.assert_array_equal(G.counter_in_type, [0, 1, 2, 3, 4, 5, 6, 7])
    np.testing.assert_array_equal(G.x, np.r_[center[:nA, 0], center[nA:, 1]])
    np.testing.assert_array_equal(G.y, np.r_[center[:nA, 1], center[nA:, 2]])
    np.testing.assert_array_equal(G.length_dir, [2, 0])
    np.testing.assert_array_equal(G.pos_side0, [center[0, 2] - 700.0, center[nA, 0] - 5.0])
    np.testing.assert_array_equal(G.pos_side1, [center[0, 2] + 700.0, center[nA, 0] + 5.0])
    np.testing.assert_array_equal(G.sipms_side1, [True, False])
    with pytest.raises(ValueError, match="counters_cfg lacks"):
        DT.sector_geom_from(geom, {k: v for k, v in cfgs.items() if k != "crs.nModulesB"})


def test_steps_placed_by_world_coordinates():
    geom, _ = DT.synthetic_crv_geometry()
    S = SY.cosmic_events(5, seed=2)
    bar = np.asarray(S["counter"])
    pos = geom.center[bar]
    W = dict(subrun=np.zeros(len(bar), np.int64), event=np.asarray(S["event"]), bar=bar, edep=np.asarray(S["edep"]),
             t0=np.asarray(S["t"]) - 0.1, t1=np.asarray(S["t"]) + 0.1)
    for a, j in (("x", 0), ("y", 1), ("z", 2)):
        off = np.asarray(S["z"]) if j == 2 else 0.0
        W[f"{a}0"] = pos[:, j] + off - (3.0 if j == 1 else 0.0)
        W[f"{a}1"] = pos[:, j] + off + (3.0 if j == 1 else 0.0)
    ev, counter, edep, z, t, order = CH.step_table(W, geom)
    e2, c2, d2, z2, t2, o2 = CH.step_table(S)
    np.testing.assert_array_equal(counter, c2)
    np.testing.assert_allclose(z, z2, rtol=0, atol=1e-9)
    np.testing.assert_allclose(t, t2, rtol=0, atol=1e-12)


def test_label_has_room_for_every_channel():
    """Reference crv/tests/test_chain.py:38-41: labels are event key * CHAN_MULT + channel; every channel of the
    detector (4 per counter) must stay below CHAN_MULT."""
    assert CF.CHAN_MULT > 1280 * 4
    assert 4 * DT.N_COUNTERS <= CF.CHAN_MULT
