"""Stage-boundary tables in the reference's dump layouts (crvhip/dumpio.py): column sets of the dumpers, round trips
through npz and parquet, the derived columns (row index in the event, member pulses, member layers) and the
dumper's branch types."""
import numpy as np
import pytest

from crvhip import chain as CH, detector as DT, dumpio as IO, synth as SY


def test_tree_columns_are_the_dumpers():
    # reference crv/art/CrvFastSim/src/CrvRecoTableDumper_module.cc:99-184 (ids, chan, then the tree's branches)
    assert IO.TREES["pulses"][:6] == ("run", "subrun", "event", "idx", "bar", "sipm")
    assert set(IO.TREES["pulses"][6:]) == {"pe", "peph", "height", "beta", "chi2", "ped", "time", "letime", "seq",
                                           "flags", "pedfromdb", "wfidx"}
    assert IO.TREES["clusters"][:4] == ("run", "subrun", "event", "idx")
    assert {"sectortype", "start", "end", "pe", "slope", "layers", "pulses", "poscalc", "avgtime", "x", "y", "z",
            "sidepe0", "sidepe1", "sidetime0", "sidetime1", "sidehits0", "sidehits1"} == set(IO.TREES["clusters"][4:])
    assert IO.TREES["calib"] == ("run", "channel", "pedestal", "pulsearea", "pulseheight", "timeoffset", "status",
                                 "roc", "feb", "febch")


@pytest.mark.parametrize("suffix", [".npz", ".parquet"])
def test_steps_round_trip(tmp_path, suffix):
    S = SY.cosmic_events(50, seed=2)
    ev, counter, edep, z, t, _ = CH.step_table(S)
    df = IO.steps_frame(ev, counter, edep, z, t, run=7)
    path = str(tmp_path / f"steps{suffix}")
    IO.write_steps(path, df)
    back = IO.read_steps(path)
    assert list(back.columns[:3]) == ["run", "subrun", "event"] and (back["run"] == 7).all()
    e2, c2, d2, z2, t2, _ = CH.step_table(back)
    for a, b in ((ev, e2), (counter, c2), (edep, d2), (z, z2), (t, t2)):
        np.testing.assert_array_equal(a, b)
    assert (back.groupby(["subrun", "event"]).size().to_numpy() == back.groupby(["subrun", "event"])["nstep"].first()
            .to_numpy()).all()


def test_read_steps_needs_the_step_columns(tmp_path):
    import pandas as pd
    p = str(tmp_path / "bad.parquet")
    pd.DataFrame({"event": [1], "bar": [3]}).to_parquet(p)
    with pytest.raises(ValueError, match="lacks columns"):
        IO.read_steps(p)


def _tables():
    """Two events (keys (0, 5) and (1, 2) as subrun << 20 | event): 3 + 2 pulses, 1 + 1 clusters."""
    k1, k2 = 5, (1 << 20) | 2
    chan = np.array([4 * 10 + 0, 4 * 26 + 1, 4 * 42 + 2, 4 * 10 + 3, 4 * 11 + 0])
    P = {"ev": np.array([k1, k1, k1, k2, k2]), "chan": chan, "wf": np.array([0, 1, 2, 3, 3]),
         "seq": np.array([0, 0, 0, 0, 1]), "pe": np.arange(5.0), "peph": np.arange(5.0) + 0.5,
         "time": np.arange(5.0) * 10, "height": np.ones(5), "beta": np.full(5, 19.0), "chi2": np.zeros(5),
         "letime": np.arange(5.0), "ped": np.full(5, 100.0), "failed": np.array([0, 1, 0, 0, 1], bool),
         "pedfromdb": np.ones(5, bool)}
    C = {"cl_event": np.array([k1, k2]), "cl_type": np.array([1, 1]), "cl_pe": np.array([3.0, 7.0]),
         "cl_start": np.zeros(2), "cl_end": np.ones(2), "cl_slope": np.zeros(2), "cl_avgtime": np.zeros(2),
         "cl_poscalc": np.array([True, False]), "cl_pos": np.arange(6.0).reshape(2, 3),
         "cl_sidepe": np.ones((2, 2)), "cl_sidetime": np.zeros((2, 2)), "cl_sidehits": np.array([[2, 1], [1, 1]]),
         "mem_cl": np.array([0, 0, 0, 1, 1]), "mem_idx": np.array([0, 1, 2, 0, 1]),
         "mem_event": np.array([k1, k1, k1, k2, k2])}
    return P, C


@pytest.mark.parametrize("suffix", [".npz", "_dir"])
def test_reco_trees_round_trip_and_derived_columns(tmp_path, suffix):
    P, C = _tables()
    G = DT.sector_geom()
    trees = IO.reco_trees({"pulses": P, "clusters": C}, DT.nominal_conditions(), G, run=3)
    p = trees["pulses"]
    np.testing.assert_array_equal(p["idx"], [0, 1, 2, 0, 1])
    np.testing.assert_array_equal(p["subrun"], [0, 0, 0, 1, 1])
    np.testing.assert_array_equal(p["event"], [5, 5, 5, 2, 2])
    np.testing.assert_array_equal(p["flags"], [0, 1, 0, 0, 1])
    np.testing.assert_array_equal(p["wfidx"].values, [0, 1, 2, 0, 0])
    c = trees["clusters"]
    # bars 10, 26, 42 lie in layers 0, 1, 2 of module 0; bars 10 and 11 both in layer 0
    assert [int(G.layer[b]) for b in (10, 26, 42, 11)] == [0, 1, 2, 0]
    np.testing.assert_array_equal(c["layers"].values, [0, 1, 2, 0])
    np.testing.assert_array_equal(c["layers"].counts, [3, 1])
    np.testing.assert_array_equal(c["pulses"].counts, [3, 2])
    np.testing.assert_array_equal(trees["events"]["npulse"], [3, 2])
    np.testing.assert_array_equal(trees["events"]["ncluster"], [1, 1])
    path = str(tmp_path / f"reco{suffix}")
    IO.write_reco(path, trees)
    for name, cols in trees.items():
        back = IO.read_reco(path, name)
        assert tuple(back) == IO.TREES[name]
        for col, v in cols.items():
            if isinstance(v, IO.Jagged):
                np.testing.assert_array_equal(back[col].values, v.values, err_msg=f"{name}.{col}")
                np.testing.assert_array_equal(back[col].counts, v.counts, err_msg=f"{name}.{col}")
            else:
                np.testing.assert_array_equal(back[col], v, err_msg=f"{name}.{col}")


def test_columns_have_the_dumpers_branch_types(tmp_path):
    # reference CrvRecoTableDumper_module.cc:59-92: unsigned ids, int counters, float PE and positions, double times
    P, C = _tables()
    trees = IO.reco_trees({"pulses": P, "clusters": C}, DT.nominal_conditions(), DT.sector_geom(), run=3)
    trees["calib"] = IO.calib_tree(DT.nominal_conditions(), run=3)
    path = str(tmp_path / "reco")
    IO.write_reco(path, trees)
    want = {("pulses", "event"): np.uint32, ("pulses", "pe"): np.float32, ("pulses", "time"): np.float64,
            ("pulses", "flags"): np.int32, ("pulses", "wfidx"): np.int32, ("clusters", "layers"): np.int32,
            ("clusters", "pe"): np.float32, ("clusters", "start"): np.float64, ("clusters", "sidehits0"): np.int32,
            ("events", "evlen"): np.float64, ("events", "npulse"): np.int32, ("calib", "pedestal"): np.float32}
    for (tree, col), dt in want.items():
        back = IO.read_reco(path, tree)[col]
        got = back.values.dtype if isinstance(back, IO.Jagged) else back.dtype
        assert got == dt, f"{tree}.{col}: {got}"
    for tree, cols in trees.items():
        assert set(IO.DTYPES[tree]) == set(IO.TREES[tree]) == set(cols)


def test_write_reco_refuses_other_columns(tmp_path):
    with pytest.raises(ValueError, match="differ"):
        IO.write_reco(str(tmp_path / "x.npz"), {"pulses": {"run": np.zeros(1)}})
