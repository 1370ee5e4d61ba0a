"""The FP32 batched chain (crv_chain_run_batched) against the FP64 chain and the oracle, through the C-ABI.

The batched chain draws the same Philox streams as the FP64 chain (which equals the oracle bit for bit) but computes
in FP32, so its tables agree with the FP64 ones statistically; the tolerances are written below. It must also be
bit-reproducible and independent of the launch shape (grid, tile, SiPM order, pulse-slot growth): each SiPM and each
reco waveform is computed by one lane, whichever lane that is.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, synth as SY  # noqa: E402
from oracle import chain as OC  # noqa: E402
from crvhip import template as TP  # noqa: E402

COUNT_RTOL = 3e-3   # pulse count, PE sum, cluster count, cluster members: FP32 chain vs FP64 chain
MATCH_DT = 1.0      # ns: a FP32 pulse matches the FP64 pulse of its channel nearest in time within this
MATCH_FRAC = 0.99   # fraction of pulses with a match
TIME_MEDIAN = 0.05  # ns: median |dt| of matched pulses
PE_MEDIAN = 2e-3    # median relative |dPE| of matched pulses


def _cfgs(**kw):
    c64 = dataclasses.replace(CF.ChainConfig(), **kw)
    return c64, dataclasses.replace(c64, precision="fp32")


def _close(a, b, rtol, slack=3):
    assert abs(a - b) <= max(slack, rtol * max(abs(a), abs(b))), (a, b)


def _compare(r32, r64, rtol=COUNT_RTOL):
    P32, P64 = r32["pulses"], r64["pulses"]
    C32, C64 = r32["clusters"], r64["clusters"]
    assert len(P64["ev"]) > 0
    _close(len(P32["ev"]), len(P64["ev"]), rtol)
    _close(P32["pe"].sum(), P64["pe"].sum(), rtol)
    _close(len(C32["cl_event"]), len(C64["cl_event"]), rtol)
    _close(len(C32["mem_idx"]), len(C64["mem_idx"]), rtol)
    # matched pulses: same (event, channel), nearest time
    k32 = P32["ev"] * 32768 + P32["chan"]
    k64 = P64["ev"] * 32768 + P64["chan"]
    o = np.lexsort((P64["time"], k64))
    k64s, t64s, pe64s = k64[o], P64["time"][o], P64["pe"][o]
    lo, hi = np.searchsorted(k64s, k32, "left"), np.searchsorted(k64s, k32, "right")
    dt = np.full(len(k32), np.inf)
    dpe = np.full(len(k32), np.inf)
    for i in np.flatnonzero(hi > lo):
        d = np.abs(t64s[lo[i]:hi[i]] - P32["time"][i])
        j = int(np.argmin(d))
        dt[i] = d[j]
        dpe[i] = abs(pe64s[lo[i] + j] - P32["pe"][i]) / max(abs(pe64s[lo[i] + j]), 1e-9)
    ok = dt < MATCH_DT
    assert ok.mean() >= MATCH_FRAC, ok.mean()
    assert np.median(dt[ok]) < TIME_MEDIAN, np.median(dt[ok])
    assert np.median(dpe[ok]) < PE_MEDIAN, np.median(dpe[ok])


@pytest.fixture(scope="module")
def cosmic():
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(3000, seed=11)
    r64 = CH.run_chain(S, cond, c64, key=7, geom=G)
    r32 = CH.run_chain(S, cond, c32, key=7, geom=G, records=True)
    return S, cond, G, c64, c32, r64, r32


def test_fp32_chain_matches_fp64_chain(cosmic):
    S, cond, G, c64, c32, r64, r32 = cosmic
    _compare(r32, r64)
    assert len(r32["pulses"]["ev"]) > 50_000


def test_fp32_chain_matches_oracle():
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(150, seed=3)
    ro = OC.run(S, cond, G, c64, 9, TP.load(c64.waveform))
    r32 = CH.run_chain(S, cond, c32, key=9, geom=G)
    Po, Co = ro["P"], ro["clusters"]
    _close(len(r32["pulses"]["ev"]), len(Po["pe"]), COUNT_RTOL)
    _close(r32["pulses"]["pe"].sum(), Po["pe"].sum(), COUNT_RTOL)
    _close(len(r32["clusters"]["cl_event"]), len(Co["cl_event"]), COUNT_RTOL)


def test_pulses_ordered_by_event_channel_time(cosmic):
    P = cosmic[6]["pulses"]
    k = P["ev"] * 32768 + P["chan"]
    assert np.all(np.diff(k) >= 0)
    same = np.diff(k) == 0
    assert np.all(np.diff(P["wf"])[same] >= 0)
    assert np.all(np.diff(P["wf"]) >= 0)


def test_launch_shape_does_not_change_results(cosmic):
    S, cond, G, c64, c32, r64, r32 = cosmic
    for batch in ({"grid": 96, "tile": 128}, {"sort_tasks": 1, "tile": 64}, {"sort_tasks": 2, "tile": 128},
                  {"sort_tasks": 3, "tile": 64}, {"pulse_slots": 1, "grid": 300}):
        r = CH.run_chain(S, cond, c32, key=7, geom=G, batch=batch)
        for k, v in r32["pulses"].items():
            np.testing.assert_array_equal(r["pulses"][k], v, err_msg=f"{batch} {k}")
        for k, v in r32["clusters"].items():
            np.testing.assert_array_equal(r["clusters"][k], v, err_msg=f"{batch} {k}")
    assert r["timing"]["pulse_slots_used"] >= 2  # the one-slot run had to grow


@pytest.mark.parametrize("sort_tasks", [1, 0])
def test_log_passes_do_not_change_results(cosmic, sort_tasks):
    """A fired-pixel log budget far below the call's log cuts the SiPMs into passes of the photon and waveform kernels
    (records and runs accumulate over the passes); the tables are those of the one-pass run."""
    S, cond, G, c64, c32, r64, r32 = cosmic
    r = CH.run_chain(S, cond, c32, key=7, geom=G, batch={"log_budget": 16 << 20, "sort_tasks": sort_tasks})
    assert r["timing"]["log_passes"] >= 3
    assert r32["timing"]["log_passes"] == 1
    for k, v in r32["pulses"].items():
        np.testing.assert_array_equal(r["pulses"][k], v, err_msg=k)
    for k, v in r32["clusters"].items():
        np.testing.assert_array_equal(r["clusters"][k], v, err_msg=k)
    assert r["timing"]["n_charges"] == r32["timing"]["n_charges"]
    assert r["timing"]["n_records"] == r32["timing"]["n_records"]


@pytest.mark.parametrize("pipeline", [2, 5])
def test_pipelined_passes_do_not_change_results(cosmic, pipeline):
    """Pipelined passes (the waveform kernel of a pass on a side stream beside the photon kernel of the next, two log
    buffers) give the tables of the one-pass run."""
    S, cond, G, c64, c32, r64, r32 = cosmic
    r = CH.run_chain(S, cond, c32, key=7, geom=G, batch={"pipeline": pipeline}, records=True)
    assert r["timing"]["log_passes"] >= pipeline
    for tab in ("pulses", "clusters", "records"):
        for k, v in r32[tab].items():
            np.testing.assert_array_equal(r[tab][k], v, err_msg=f"{tab} {k}")


def test_log_budget_below_one_group_is_an_error(cosmic):
    from crvhip import _lib as L
    S, cond, G, c64, c32, r64, r32 = cosmic
    with pytest.raises(L.CrvError, match="exceeds the log budget"):
        CH.run_chain(S, cond, c32, key=7, geom=G, batch={"log_budget": 64})
    with pytest.raises(L.CrvError, match="sort_tasks must be"):
        CH.run_chain(S, cond, c32, key=7, geom=G, batch={"sort_tasks": 4})


def test_counter_beyond_the_conditions_is_an_error():
    """A step on a counter whose channels the conditions do not hold is an error in both chains (the reference
    indexes its status table with the channel, chain.py:200, and fails there)."""
    from crvhip import _lib as L
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = {k: np.array(v) for k, v in SY.cosmic_events(20, seed=3).items()}
    S["counter"][len(S["counter"]) // 2] = len(cond.status) // 4 + 2
    for cfg in (c64, c32):
        with pytest.raises(L.CrvError, match="beyond the uploaded conditions"):
            CH.run_chain(S, cond, cfg, key=1, geom=G)
    # the context stays usable: the valid steps run afterwards
    S = SY.cosmic_events(20, seed=3)
    assert len(CH.run_chain(S, cond, c32, key=1, geom=G)["pulses"]["ev"]) > 0


def test_records_match_fp64_records(cosmic):
    S, cond, G, c64, c32, r64, r32 = cosmic
    R = r32["records"]
    assert len(R["label"]) > 10_000
    assert np.all(R["adc"] <= 4095)
    Q = CH.charges(S, None, G, cond, c64, None, key=7)
    D = CH.digis(Q, cond, c64, None, key=7)
    k64 = D["record_label"] * 2 ** 20 + D["record_tick"]
    k32 = R["label"] * 2 ** 20 + R["tick"]
    _close(len(k32), len(k64), 0.01, slack=20)
    both, i32, i64 = np.intersect1d(k32, k64, return_indices=True)
    assert len(both) >= 0.99 * len(k64)
    same = R["adc"][i32].astype(np.int32) == D["record_adc"][i64].astype(np.int32)
    assert same.mean() > 0.99, same.mean()  # same noise draws; FP32 sums move a word by one now and then


def test_empty_and_tiny_inputs():
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(1, seed=0)
    empty = {k: v[:0] for k, v in S.items()}
    r = CH.run_chain(empty, cond, c32, key=1, geom=G)
    assert len(r["pulses"]["ev"]) == 0 and len(r["clusters"]["cl_event"]) == 0
    S = SY.cosmic_events(5, seed=1)
    _r64, _r32 = CH.run_chain(S, cond, c64, key=2, geom=G), CH.run_chain(S, cond, c32, key=2, geom=G)
    _close(len(_r32["pulses"]["ev"]), len(_r64["pulses"]["ev"]), 0.05)


def test_heavy_cross_talk_and_dead_channels():
    c64, c32 = _cfgs(sipm=dataclasses.replace(CF.SiPMConfig(), cross_talk=0.2))
    cond = DT.nominal_conditions()
    st = np.array(cond.status)
    rng = np.random.default_rng(4)
    st[rng.choice(len(st), 300, replace=False)] = rng.choice([1, 2, 4], 300)
    cond = dataclasses.replace(cond, status=st)
    G = DT.sector_geom()
    S = SY.cosmic_events(1500, seed=21)
    _compare(CH.run_chain(S, cond, c32, key=5, geom=G), CH.run_chain(S, cond, c64, key=5, geom=G), rtol=5e-3)


def test_uneven_showers():
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.uneven_events(300, seed=8)
    r32 = CH.run_chain(S, cond, c32, key=3, geom=G)
    _compare(r32, CH.run_chain(S, cond, c64, key=3, geom=G), rtol=5e-3)
    # counters of many steps (merged photon streams) and events for the block finder: the same tables in every
    # queue order and tile
    for batch in ({"sort_tasks": 2, "tile": 64}, {"sort_tasks": 3, "tile": 128}, {"sort_tasks": 1, "tile": 256}):
        r = CH.run_chain(S, cond, c32, key=3, geom=G, batch=batch)
        for tab in ("pulses", "clusters"):
            for k, v in r32[tab].items():
                np.testing.assert_array_equal(r[tab][k], v, err_msg=f"{batch} {tab} {k}")


def test_overlay_events():
    """The overlay workload (~2,800 steps and ~5,500 pulses per event, pile-up in the fits): FP32 against FP64."""
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.overlay_events(40, seed=9)
    r64 = CH.run_chain(S, cond, c64, key=3, geom=G)
    r32 = CH.run_chain(S, cond, c32, key=3, geom=G)
    assert len(r64["pulses"]["ev"]) > 100_000
    _compare(r32, r64)


def test_autotuner_picks_a_candidate_and_results_do_not_change(cosmic, tmp_path):
    """The work-queue autotuner times every candidate on device-resident steps and returns the fastest; the chain
    run with its choice gives the same tables as the default launch."""
    from crvhip import _lib as L, autotune as AT
    S, cond, G, c64, c32, r64, r32 = cosmic
    ev, counter, edep, z, t, _ = CH.step_table(S)
    ds = CH.DeviceSteps(ev, counter, edep, z, t)
    ctx = CH.context(c32, cond, G)
    ctx.set_params(c32, 7, ds.evkey_hi, 32)
    cache = str(tmp_path / "tune.json")
    chosen, rep = AT.tune(ctx, ds.struct(), 3000, "test-device", tiles=(64, 128, 256), sorts=(1, 0), repeats=1,
                          cache_file=cache)
    assert not rep["cached"] and len(rep["candidates"]) == 6
    assert all(len(c["ms"]) == 1 and c["ms"][0] > 0 for c in rep["candidates"])
    assert chosen in [c["opts"] for c in rep["candidates"]]
    again, rep2 = AT.tune(ctx, ds.struct(), 3000, "test-device", cache_file=cache)
    assert rep2["cached"] and again == chosen
    r = CH.run_chain(S, cond, c32, key=7, geom=G, batch=chosen)
    for k, v in r32["pulses"].items():
        np.testing.assert_array_equal(r["pulses"][k], v, err_msg=k)
    for k, v in r32["clusters"].items():
        np.testing.assert_array_equal(r["clusters"][k], v, err_msg=k)


def test_pipelined_passes_on_small_and_skewed_calls():
    """More pipelined passes than 64-SiPM groups (a few events), and a longest-first uneven queue whose first group
    holds more than its share of the log: the passes hold at least the largest group, results are the one-pass run's."""
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    for S, batch in ((SY.cosmic_events(1, seed=4), {}), (SY.cosmic_events(5, seed=4), {}),
                     (SY.uneven_events(200, seed=5), {"sort_tasks": 2}), (SY.uneven_events(200, seed=5),
                                                                           {"sort_tasks": 1})):
        r0 = CH.run_chain(S, cond, c32, key=4, geom=G, batch=batch)
        r = CH.run_chain(S, cond, c32, key=4, geom=G, batch={**batch, "pipeline": 5})
        for tab in ("pulses", "clusters"):
            for k, v in r0[tab].items():
                np.testing.assert_array_equal(r[tab][k], v, err_msg=f"{batch} {tab} {k}")


def test_trim_returns_the_workspace_to_the_caller():
    """After a call the library keeps its workspace for the next one; a caller allocation that does not fit next to
    it trims the workspace and succeeds, and the next call gives the same tables."""
    import torch
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SY.cosmic_events(30_000, seed=6)
    r0 = CH.run_chain(S, cond, c32, key=5, geom=G)
    ctx = CH.context(c32, cond, G)
    ws = ctx.workspace_bytes()
    assert ws > 256 << 20, ws
    torch.cuda.empty_cache()
    free = torch.cuda.mem_get_info()[0]
    want = free + ws // 2
    with pytest.raises(torch.OutOfMemoryError):
        torch.empty(want, dtype=torch.uint8, device="cuda")
    t = CH._empty(want, torch.uint8)
    assert ctx.workspace_bytes() == 0
    del t
    torch.cuda.empty_cache()
    r = CH.run_chain(S, cond, c32, key=5, geom=G)
    for tab in ("pulses", "clusters"):
        for k, v in r0[tab].items():
            np.testing.assert_array_equal(r[tab][k], v, err_msg=f"{tab} {k}")
    assert CH.trim() > 0 and ctx.workspace_bytes() == 0


def _twice_hit(n, seed):
    """Cosmic events whose counters are each hit again 700 ns later: two kept charge clusters per SiPM."""
    S = SY.cosmic_events(n, seed=seed)
    S2 = {k: np.concatenate([v, v]) for k, v in S.items()}
    S2["t"][len(S["t"]):] += 700.0
    o = np.argsort(S2["event"], kind="stable")
    return {k: v[o] for k, v in S2.items()}, len(S["t"])


def test_pool_growth_does_not_change_results():
    """Two kept clusters per SiPM: the first call on a fresh context outgrows the descriptor, record and run pools it
    starts with (sized for one cluster per SiPM), grows them and runs the photon and waveform kernels again. The
    tables equal a second call's, whose pools are large enough from the start, a pipelined fresh context's, and the
    FP64 chain's within the tolerances above."""
    c64, c32 = _cfgs()
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S, nsteps = _twice_hit(600, seed=31)
    r1 = CH.run_chain(S, cond, c32, key=3, geom=G, records=True)
    assert len(r1["records"]["label"]) > 4 * nsteps + 1024  # more than the starting pool (SiPMs + 1024): it grew
    r2 = CH.run_chain(S, cond, c32, key=3, geom=G, records=True)
    r3 = CH.run_chain(S, cond, dataclasses.replace(c32), key=3, geom=G, batch={"pipeline": 2}, records=True)
    for r in (r2, r3):
        for tab in ("pulses", "clusters", "records"):
            for k, v in r1[tab].items():
                np.testing.assert_array_equal(r[tab][k], v, err_msg=f"{tab} {k}")
    _compare(r1, CH.run_chain(S, cond, c64, key=3, geom=G))
