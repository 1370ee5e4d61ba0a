"""The FP64 plain chain on the named workloads: the CPU parity run (1,000 cosmic events, seed 0) bit-exact against
the oracle through clusters, the uneven mixture at default settings, and 100,000 events through size-independent
properties (one call equals four calls over contiguous event ranges; FP32 counts within 0.3 %)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, synth as SY, template as TP  # noqa: E402
from oracle import chain as OC  # noqa: E402

INT_PULSE = ("wf", "seq", "failed", "zerondf", "pedfromdb")
F_PULSE = ("pe", "peph", "time", "height", "beta", "chi2", "letime", "ped", "fit_a", "fit_mu", "fit_b")
INT_CLUSTER = ("cl_event", "cl_type", "cl_poscalc", "cl_sidehits", "mem_cl", "mem_idx", "mem_event")
F_CLUSTER = ("cl_pe", "cl_start", "cl_end", "cl_slope", "cl_avgtime", "cl_pos", "cl_sidepe", "cl_sidetime")


def _against_oracle(S, key):
    cfg = CF.ChainConfig()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    r = CH.run_chain(S, cond, cfg, key=key, geom=G)
    ro = OC.run(S, cond, G, cfg, key, TP.load(cfg.waveform))
    P, Po = r["pulses"], ro["P"]
    o = np.lexsort((P["seq"], P["wf"]))
    oo = np.lexsort((Po["seq"], Po["wf"]))
    assert len(o) == len(oo) > 0
    for k in INT_PULSE:
        np.testing.assert_array_equal(np.asarray(P[k])[o].astype(np.int64), np.asarray(Po[k])[oo].astype(np.int64),
                                      err_msg=k)
    for k in F_PULSE:
        np.testing.assert_allclose(P[k][o], Po[{"fit_a": "fit_A"}.get(k, k)][oo], rtol=1e-12, atol=0, err_msg=k)
    np.testing.assert_array_equal(P["ev"][o], ro["W"]["ev"][Po["wf"][oo]])
    C, Co = r["clusters"], ro["clusters"]
    assert len(Co["cl_event"]) > 0
    for k in INT_CLUSTER:
        np.testing.assert_array_equal(np.asarray(C[k]).astype(np.int64), np.asarray(Co[k]).astype(np.int64),
                                      err_msg=k)
    for k in F_CLUSTER:
        np.testing.assert_allclose(np.asarray(C[k], float), np.asarray(Co[k], float), rtol=1e-12, atol=1e-9,
                                   err_msg=k)
    return r


def test_cpu_parity_run_bit_exact_through_clusters():
    """1,000 cosmic events of seed 0, Philox seed 0: every pulse and cluster of the FP64 chain equals the oracle's."""
    r = _against_oracle(SY.cosmic_events(1000, seed=0), 0)
    assert len(r["clusters"]["cl_event"]) > 2500


def test_uneven_mixture_bit_exact_through_clusters():
    """300 events of the uneven mixture (cosmic, 9 % showers over 8 modules, 1 % dense showers) at default SiPM
    settings."""
    r = _against_oracle(SY.uneven_events(300, seed=2), 3)
    assert r["timing"]["n_pulses"] > 25_000


def _split(S, parts):
    ev = np.asarray(S["event"])
    cuts = np.searchsorted(ev, np.quantile(np.unique(ev), np.linspace(0, 1, parts + 1)[1:-1]))
    b = np.r_[0, cuts, len(ev)]
    return [{k: np.asarray(v)[b[i]:b[i + 1]] for k, v in S.items()} for i in range(parts)]


def test_fp64_chain_at_100k_events_one_call_equals_four():
    cfg = CF.ChainConfig()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(100_000, seed=0)
    r = CH.run_chain(S, cond, cfg, key=0, geom=G)
    parts = [CH.run_chain(s, cond, cfg, key=0, geom=G) for s in _split(S, 4)]
    P = r["pulses"]
    assert np.all(P["pe"] > 0) and np.all(np.isfinite(P["time"]))   # every pulse written (no stale workspace)
    wf0 = np.cumsum([0] + [int(p["timing"]["n_waveforms"]) for p in parts[:-1]])
    for k in P:
        cat = np.concatenate([p["pulses"][k] + (wf0[i] if k == "wf" else 0) for i, p in enumerate(parts)])
        np.testing.assert_array_equal(P[k], cat, err_msg=k)
    C = r["clusters"]
    n0 = np.cumsum([0] + [len(p["clusters"]["cl_event"]) for p in parts[:-1]])
    for k in C:
        cat = np.concatenate([p["clusters"][k] + (n0[i] if k == "mem_cl" else 0) for i, p in enumerate(parts)])
        np.testing.assert_array_equal(C[k], cat, err_msg=k)
    assert len(C["cl_event"]) > 250_000
    r32 = CH.run_chain(S, cond, dataclasses.replace(cfg, precision="fp32"), key=0, geom=G)
    for a, b in ((len(r32["pulses"]["pe"]), len(P["pe"])), (r32["pulses"]["pe"].sum(), P["pe"].sum()),
                 (len(r32["clusters"]["cl_event"]), len(C["cl_event"])),
                 (len(r32["clusters"]["mem_idx"]), len(C["mem_idx"]))):
        assert abs(a - b) <= 3e-3 * b, (a, b)
