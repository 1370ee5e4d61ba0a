"""The FP32 batched chain captured into one HIP graph (crv_graph_*) against the same chain run uncaptured: every
replay gives the uncaptured tables bit for bit, with eager runs between replays, on cosmic and uneven batches of
other seeds than the sample the graph was sized on; calls beyond the plan's capacities run uncaptured (same tables)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import chain as CH, config as CF, detector as DT, synth as SY  # noqa: E402


def _same(a, b, what):
    for tab in ("pulses", "clusters"):
        for k, v in b[tab].items():
            np.testing.assert_array_equal(a[tab][k], v, err_msg=f"{what} {tab} {k}")


@pytest.fixture(scope="module")
def setup():
    cfg = dataclasses.replace(CF.ChainConfig(), precision="fp32")
    return cfg, DT.nominal_conditions(), DT.sector_geom()


def test_graph_replays_equal_the_uncaptured_chain_on_cosmic_batches(setup):
    cfg, cond, G = setup
    batch = {"tile": 128, "sort_tasks": 3}
    sample = SY.cosmic_events(3000, seed=31)
    g = CH.ChainGraph(sample, cond, cfg, key=7, geom=G, batch=batch, max_steps=int(len(sample["event"]) * 1.5))
    others = [SY.cosmic_events(3000, seed=s) for s in (32, 33)] + [SY.cosmic_events(1200, seed=34)]
    eager = [CH.run_chain(S, cond, cfg, key=7, geom=G, batch=batch) for S in [sample] + others]
    for rep in range(3):   # replays, eager work of the context between them
        for S, ref in zip([sample] + others, eager):
            r = g.run(S)
            _same(r, ref, f"replay {rep}")
            assert r["timing"]["n_pulses"] == len(ref["pulses"]["pe"]) > 0
            assert r["timing"]["k_sipm_ms"] > 0 and r["timing"]["k_pulse_ms"] > 0   # kernel times of the replay
        CH.run_chain(others[rep % len(others)], cond, cfg, key=7, geom=G, batch=batch)
    st = g.graph.stats()
    assert st["fallbacks"] == 0 and st["launches"] == 12
    g.close()


def test_graph_on_uneven_batches_and_fallbacks(setup):
    cfg, cond, G = setup
    batch = {"tile": 64, "sort_tasks": 2}
    sample = SY.uneven_events(400, seed=41)
    g = CH.ChainGraph(sample, cond, cfg, key=3, geom=G, batch=batch, max_steps=int(len(sample["event"]) * 1.3))
    for seed in (42, 43, 44):
        S = SY.uneven_events(400, seed=seed)
        r = g.run(S)
        _same(r, CH.run_chain(S, cond, cfg, key=3, geom=G, batch=batch), f"uneven seed {seed}")
    # a call of more steps than the graph holds runs uncaptured, with the same tables
    big = SY.uneven_events(1000, seed=45)
    r = g.run(big)
    assert r["graph"]["capacities"]["last_fell_back"] == 1
    _same(r, CH.run_chain(big, cond, cfg, key=3, geom=G, batch=batch), "beyond capacity")
    # back within capacity: captured again
    S = SY.uneven_events(400, seed=46)
    r = g.run(S)
    assert r["graph"]["capacities"]["last_fell_back"] == 0
    _same(r, CH.run_chain(S, cond, cfg, key=3, geom=G, batch=batch), "after fallback")
    g.close()


def test_graph_empty_and_tiny_calls(setup):
    cfg, cond, G = setup
    sample = SY.cosmic_events(200, seed=51)
    g = CH.ChainGraph(sample, cond, cfg, key=2, geom=G, batch={"tile": 64, "sort_tasks": 1})
    empty = {k: v[:0] for k, v in SY.cosmic_events(1, seed=0).items()}
    r = g.run(empty)
    assert len(r["pulses"]["pe"]) == 0 and len(r["clusters"]["cl_event"]) == 0
    S = SY.cosmic_events(3, seed=52)
    _same(g.run(S), CH.run_chain(S, cond, cfg, key=2, geom=G, batch={"tile": 64, "sort_tasks": 1}), "tiny")
    g.close()
