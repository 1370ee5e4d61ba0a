"""Onspill timing contracts of crvhip.onspill on the CPU: event population and full-ID lookup, the step prefilter,
arrival folding, the bunch correction, configuration checks and the physical sensor mask. Re-expresses the
reference's crv/tests/test_onspill.py cases that need no device (its cases that spy on the optics and the SiPM
stage run against the HIP path in tests/test_gpu_onspill.py)."""
from dataclasses import replace
from types import SimpleNamespace

import numpy as np
import pytest

from crvhip import chain as CH, config as CF, onspill as O, readout as RO
from oracle import onspill as OO


def timing(n=1, **kw):
    args = dict(keys=np.column_stack((np.ones(n, int), np.zeros(n, int), np.arange(n))),
                mc_pbtime=np.zeros(n), reco_pbtime=np.zeros(n), event_length=np.full(n, 1675.))
    return O.EventTiming(**{**args, **kw})


def conditions():
    z = np.zeros(8)
    return CH.Conditions(z + 100, z + 400, z + 11, z + 7, z.astype(int), z.astype(int), z.astype(int), z,
                         readout_mask=np.ones(8, bool))


def config(**kw):
    return replace(CF.ChainConfig(), window=(200, 1500), sipm=replace(CF.SiPMConfig(), thermal_rate=1e-4), **kw)


def steps(t0=(300.,), t1=None, keys=None):
    n = len(t0)
    keys = np.tile([1, 0, 0], (n, 1)) if keys is None else np.asarray(keys)
    t1 = t0 if t1 is None else t1
    return {"run": keys[:, 0] if n else np.zeros(0, int), "subrun": keys[:, 1] if n else np.zeros(0, int),
            "event": keys[:, 2] if n else np.zeros(0, int), "counter": np.zeros(n, int), "edep": np.ones(n),
            "z": np.zeros(n), "t0": np.asarray(t0, float), "t1": np.asarray(t1, float)}


def test_exact_prefilter_boundaries_crossings_original_intervals():
    """Reference test_onspill.py:64-77: below the prestart margin, above the window, the excluded crossing go;
    equality at either bound stays; a step crossing the period boundary stays."""
    a = np.array([99., 100., 1500., 1501., 1600., 1600., 1690., -5., 1795., 99.])
    b = np.array([99., 100., 1500., 1501., 50., 100., 1805., 110., 1795., 100.])
    rows, finite = O.prefilter(np.column_stack([a, b]), np.zeros(len(a)), 1695., (200, 1500))
    np.testing.assert_array_equal(rows, [1, 2, 5, 6, 7, 8, 9])
    assert finite.all()
    np.testing.assert_array_equal(OO.gate(a, b, np.zeros(len(a)), 1695., 200, 1500), rows)
    rows, finite = O.prefilter(np.array([[np.nan, 300.], [300., 310.]]), np.zeros(2), 1695., (200, 1500))
    np.testing.assert_array_equal(rows, [1])
    np.testing.assert_array_equal(finite, [False, True])


def test_arrival_fold_negative_and_prestart_retained():
    """Reference test_onspill.py:80-85: a shift of 110 folds arrivals into [0, 1695), negatives moved up."""
    t = np.array([-1805., -110., -105., 1395., 1400., 1585., 3280.])
    np.testing.assert_array_equal(O.fold(t + 110., 1695.), [0, 0, 5, 1505, 1510, 0, 0])
    np.testing.assert_array_equal(OO.fold(t + 110., 1695.), [0, 0, 5, 1505, 1510, 0, 0])


def test_same_shift_invariant_and_true_reco_separation():
    """Reference test_onspill.py:95-105: steps 1699 ns later in events 1699 ns earlier select and fold the same;
    the reconstructed bunch correction is -reco_pbtime."""
    ends = np.array([[290., 310.], [301., 340.]])
    r1, _ = O.prefilter(ends, np.full(2, 10.), 1695., (200, 1500))
    r2, _ = O.prefilter(ends + 1699, np.full(2, 10. - 1699), 1695., (200, 1500))
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(O.fold(ends + 10., 1695.), O.fold(ends + 1699 + (10. - 1699), 1695.))
    tm = timing(mc_pbtime=np.array([10.]), reco_pbtime=np.array([80.]))
    W = O.reco_timing(dict(ev=np.array([0, 0])), tm)
    np.testing.assert_array_equal(W["event_t0"], [-80, -80])
    with pytest.raises(ValueError, match="dense timing row"):
        O.reco_timing(dict(ev=np.array([1])), tm)


def test_uint64_full_ids_and_signed_lookup():
    """Reference test_onspill.py:125-133."""
    keys = np.array([[2 ** 64 - 1, 4, 5], [1, 0, 0]], dtype=np.uint64)
    tm = timing(2, keys=keys)
    np.testing.assert_array_equal(tm.lookup(keys[::-1]), [1, 0])
    np.testing.assert_array_equal(tm.lookup(np.array([[1, 0, 0]], np.int64)), [1])
    with pytest.raises(ValueError, match="nonnegative"):
        tm.lookup(np.array([[-1, 4, 5]], np.int64))
    with pytest.raises(ValueError, match="absent"):
        timing().lookup(keys[:1])
    big = np.array([[2 ** 61, 2 ** 42, 2 ** 40], [2 ** 61 + 1, 2 ** 42, 2 ** 40], [3, 4, 5]], np.int64)
    np.testing.assert_array_equal(timing(3, keys=big).lookup(big[[1, 0, 1, 0]]), [1, 0, 1, 0])


@pytest.mark.parametrize("kw,match", [(dict(keys=np.array([[1, 0, 0], [1, 0, 0]])), "unique"),
                                      (dict(mc_pbtime=np.array([np.nan])), "mc_pbtime"),
                                      (dict(event_length=np.array([0.])), "event_length"),
                                      (dict(period=0.), "period"), (dict(keys=np.array([[1, 0]])), "Nx3")])
def test_invalid_timing(kw, match):
    n = len(kw["keys"]) if "keys" in kw else 1
    base = dict(keys=np.column_stack((np.ones(n, int), np.zeros(n, int), np.zeros(n, int))) if "keys" not in kw
                else kw["keys"], mc_pbtime=np.zeros(n), reco_pbtime=np.zeros(n), event_length=np.full(n, 1675.))
    with pytest.raises(ValueError, match=match):
        O.EventTiming(**{**base, **kw})


@pytest.mark.parametrize("t0,n_events", [((300.,), 2), ((), 2), ((), 0)])
@pytest.mark.parametrize("thermal_rate", [0., 1e-4])
def test_missing_readout_mask_rejected_before_response(t0, n_events, thermal_rate):
    """Reference test_onspill.py:136-147: no device work, conditions untouched."""
    cond = replace(conditions(), readout_mask=None)
    cfg = config()
    cfg = replace(cfg, sipm=replace(cfg.sipm, thermal_rate=thermal_rate))
    st = steps(t0)
    with pytest.raises(ValueError, match="require cond.readout_mask"):
        O.charges(st, np.full(len(t0), 13), SimpleNamespace(nbars=2), cond, cfg, 123, timing=timing(n_events))
    assert cond.readout_mask is None
    np.testing.assert_array_equal(cond.status, np.zeros(8, int))


@pytest.mark.parametrize("counters", [{}, {"crs.sipmsAtSide0A": True},
                                      {"crs.sipmsAtSide0A": True, "crs.sipmsAtSide1A": 1}])
def test_run_missing_mask_never_guesses_side_presence(counters):
    """Reference test_onspill.py:196-205."""
    geom = SimpleNamespace(nbars=2, sector_names=["A"], sector=np.zeros(2, int))
    cond = replace(conditions(), readout_mask=None)
    with pytest.raises(ValueError, match="readout counters"):
        O.run(steps([]), np.empty(0, int), geom, cond, config(), 123, timing=timing(0), counters_cfg=counters)
    assert cond.readout_mask is None


def test_physical_sides_mask():
    """Reference test_onspill.py:150-193's sensor mask: side 0 of sector A and side 1 of sector B are present."""
    geom = SimpleNamespace(nbars=2, sector_names=["A", "B"], sector=np.array([0, 1]))
    counters = {"crs.sectorNames": ["A", "B"], "crs.sipmsAtSide0A": True, "crs.sipmsAtSide1A": False,
                "crs.sipmsAtSide0B": False, "crs.sipmsAtSide1B": True}
    mask = np.array([True, False, True, False, False, True, False, True])
    np.testing.assert_array_equal(RO.sensor_mask(geom, counters), mask)
    cond = conditions()
    cond.status[:] = [0, 0, 2, 0, 0, 0, 0, 4]
    masked = RO.apply(replace(cond, readout_mask=None), geom, counters)
    np.testing.assert_array_equal(masked.readout_mask, mask)
    np.testing.assert_array_equal(cond.status, [0, 0, 2, 0, 0, 0, 0, 4])
    eff = RO.effective_status(masked)
    np.testing.assert_array_equal(np.flatnonzero((eff & 0b101) == 0), [0, 2, 5])
    with pytest.raises(ValueError, match="disagree"):
        RO.sensor_mask(geom, {**counters, "crs.sectorNames": ["B", "A"]})


def test_empty_population_no_device_work():
    """Reference test_onspill.py:241-247: no steps and no events give empty tables without any device call; steps
    of events outside the population are an error."""
    Q = O.charges(steps([]), np.empty(0, int), SimpleNamespace(nbars=2), conditions(), config(), 123,
                  timing=timing(0))
    assert all(len(v) == 0 for v in Q.values())
    assert Q["chan"].dtype.kind == Q["photon"].dtype.kind == "i"
    with pytest.raises(ValueError, match="absent"):
        O.charges(steps(), np.array([13]), SimpleNamespace(nbars=2), conditions(), config(), 123, timing=timing(0))


@pytest.mark.parametrize("change,match", [(dict(window=(50, 1500)), "prestart"), (dict(window=(200, 1800)), "period"),
                                          ("traps", "trap afterpulses"), ("p1", "p1"), ("tabs", "tabs"),
                                          ("pdg", "pdg")])
def test_invalid_configuration(change, match):
    cfg, kw, pdg = config(), {}, np.array([13])
    if change == "traps":
        cfg = replace(cfg, sipm=replace(cfg.sipm, trap0_prob=0.1))
    elif change == "p1":
        cfg = replace(cfg, sipm=replace(cfg.sipm, p1=1.5))
    elif change == "tabs":
        kw = dict(tabs={})
    elif change == "pdg":
        pdg = np.array([13.0])
    else:
        cfg = replace(cfg, **change)
    with pytest.raises(ValueError, match=match):
        O.charges(steps(), pdg, SimpleNamespace(nbars=2), conditions(), cfg, 123, timing=timing(), **kw)
    with pytest.raises(ValueError, match="event_length"):
        O.charges(steps(), np.array([13]), SimpleNamespace(nbars=2), conditions(), config(), 123,
                  timing=timing(event_length=np.array([1400.])))
