"""The chain driven from a step file in the CrvStepTableDumper layout and written back in the CrvRecoTableDumper
layout (crvhip/dumpio.py), through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, dumpio as IO, synth as SY  # noqa: E402


@pytest.mark.parametrize("suffix", [".parquet", ".npz"])
def test_step_file_to_crvdump(tmp_path, suffix):
    cfg, cond, G = CF.ChainConfig(), DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(300, seed=8)
    ev, counter, edep, z, t, _ = CH.step_table(S)
    steps_path = str(tmp_path / f"steps{suffix}")
    IO.write_steps(steps_path, IO.steps_frame(ev, counter, edep, z, t, run=12))
    steps = IO.read_steps(steps_path)
    direct = CH.run_chain(S, cond, cfg, key=4, geom=G)
    out = CH.run_chain(steps, cond, cfg, key=4, geom=G)
    for tab in ("pulses", "clusters"):
        for k, v in direct[tab].items():
            np.testing.assert_array_equal(out[tab][k], v, err_msg=f"{tab}.{k}")
    trees = IO.reco_trees(out, cond, G, run=12)
    reco_path = str(tmp_path / ("reco.npz" if suffix == ".npz" else "reco"))
    IO.write_reco(reco_path, trees)
    P, C = IO.read_reco(reco_path, "pulses"), IO.read_reco(reco_path, "clusters")
    assert len(P["idx"]) == len(out["pulses"]["ev"]) > 1000 and len(C["idx"]) == len(out["clusters"]["cl_event"])
    np.testing.assert_array_equal(P["time"], out["pulses"]["time"])
    np.testing.assert_array_equal(C["pe"], out["clusters"]["cl_pe"].astype(np.float32))  # a float branch
    # a cluster's members, looked up by (event, idx) in the pulse tree, are the chain's member pulses
    pkey = P["subrun"].astype(np.int64) << 20 | P["event"]
    ckey = np.repeat(C["subrun"].astype(np.int64) << 20 | C["event"], C["pulses"].counts)
    rows = np.searchsorted(pkey, ckey) + C["pulses"].values
    np.testing.assert_array_equal(pkey[rows], ckey)
    np.testing.assert_array_equal(P["idx"][rows], C["pulses"].values)
    nlay = C["layers"].counts  # kept hits pass a three-of-four-layer combination
    assert ((nlay >= 3) & (nlay <= 4)).all()
    lay = C["layers"].values
    first = np.repeat(np.r_[0, np.cumsum(nlay)[:-1]], nlay)
    assert ((lay >= 0) & (lay < 4)).all() and (np.diff(lay)[np.diff(first) == 0] > 0).all()  # sorted, distinct
    E = IO.read_reco(reco_path, "events")
    assert E["npulse"].sum() == len(P["idx"]) and E["ncluster"].sum() == len(C["idx"])
    assert (E["run"] == 12).all()
