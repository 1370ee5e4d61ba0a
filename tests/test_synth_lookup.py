"""The lookup-optics workload of bench.py (crvhip.synth_lookup) on the CPU: its table is a valid v6 table of the
synthetic detector, and through the oracle's restatement of the reference's lookup optics it gives about the
photoelectrons of the closed-form optics on the same cosmic steps (the property the workload is built for)."""
import numpy as np

from crvhip import chain as CH, config as CF, detector as DT, lookup as LK, photons as PH, synth as SY
from crvhip import synth_lookup as SL
from oracle import lookup_optics as LO, response as RS


def test_table_is_a_v6_table_of_the_synthetic_counters(tmp_path):
    T = SL.table(nz=24)
    LK.write(T, tmp_path / "LookupTable_synthetic")
    U = LK.load("synthetic", str(tmp_path))
    assert U.half_length == DT.HALF_LENGTH and U.reflector == 0
    for k in LK.TABLES:
        assert np.array_equal(U.tables[k].prob, T.tables[k].prob)
        assert np.array_equal(U.tables[k].delay_sum, T.tables[k].delay_sum)
    inv_v, att = PH.table_propagation(T)
    assert abs(1.0 / inv_v - 160.0) < 8.0        # delays peak at d / 160 mm/ns
    assert 3000.0 < att < 6000.0                  # the long attenuation component dominates the fit


def test_photons_match_the_closed_form_optics_on_average():
    S = SY.cosmic_events(40, seed=3)
    cols, pdg = SL.crv_steps(S)
    T = SL.table(nz=102)
    G, _ = DT.synthetic_crv_geometry()
    bars, tl = PH.bar_arrays(G, SL.response_config(), {T.name: T})
    cond = DT.nominal_conditions()
    m, q, _ = PH.particle_props(pdg)
    st = {"evkey": CH.event_key(cols), "bar": cols["bar"], "t0": cols["t0"], "t1": cols["t1"],
          "pos0": np.column_stack([cols[f"{a}0"] for a in "xyz"]),
          "pos1": np.column_stack([cols[f"{a}1"] for a in "xyz"]), "edep": cols["edep"], "len": cols["len"],
          "p0": np.abs(cols["py0"]), "p1": cols["p1"], "mass": m, "charge": q}
    props = [LO.table_propagation(X) for X in tl]
    lk = LO.photons(st, bars, tl, props, cond.deviation, DT.effective_status(cond), 1.0, 1.0, 5)
    cf = RS.photons({k: S[k] for k in ("evkey", "counter", "edep", "z", "t")}, DT.effective_status(cond),
                    np.asarray(cond.deviation) + 1.0, CF.ChainConfig().optics, 5)
    n_lk, n_cf = len(lk["time"]), len(cf["time"])
    assert n_cf > 50_000
    assert abs(n_lk - n_cf) < 0.05 * n_cf
