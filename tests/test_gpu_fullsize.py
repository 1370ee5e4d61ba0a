"""The FP32 batched chain at its full size, 1,000,000 cosmic events generated on the device, checked through properties
that do not depend on the size (the oracle cannot run this many events):

* the tables of one call equal those of the same events run in four calls of contiguous event ranges, every column
  bit for bit, the member offsets shifted by the earlier calls' member counts (every draw is keyed by event, channel,
  step and photon; events share no pulse and no cluster; a member is a pulse index within its event);
* pulses come in (event, channel, time) order;
* every cluster member is a pulse of the cluster's own event, the member offsets rise from 0 to the member count.

The comparisons run on the device; nothing of the million-event tables is copied to the host.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

N_EVENTS = 1_000_000
PULSE_COLS = ("seq", "ev", "chan", "pe", "peph", "time", "height", "beta", "chi2", "letime", "ped", "fit_a", "fit_mu",
              "fit_b", "failed", "zerondf", "pedfromdb")
CLUSTER_COLS = ("cl_event", "cl_type", "cl_poscalc", "cl_sidehits", "cl_pe", "cl_start", "cl_end", "cl_slope",
                "cl_avgtime", "cl_pos", "cl_sidepe", "cl_sidetime")


def _run(torch, L, CH, ctx, st):
    """The chain over the steps `st` (a crv_steps): device tables of exactly the needed sizes."""
    opts = L.BatchOpts()
    pb, ps = CH._pulse_bufs(0)
    cb, cs = CH._cluster_bufs(0, 0)
    ps.capacity = cs.capacity = cs.mem_capacity = 0
    tm = L.Timing()
    rc = ctx.lib.crv_chain_run_batched(ctx.h, ctypes.byref(st), ctypes.byref(opts), ctypes.byref(ps),
                                       ctypes.byref(cs), ctypes.byref(tm), None)
    assert rc in (0, L.CRV_E_CAPACITY), rc
    pb, ps = CH._pulse_bufs(max(ps.n, 1))
    cb, cs = CH._cluster_bufs(max(cs.n, 1), max(cs.n_mem, 1))
    ctx.check(ctx.lib.crv_chain_fetch(ctx.h, ctypes.byref(ps), ctypes.byref(cs), None))
    torch.cuda.synchronize()
    return pb, ps.n, cb, cs.n, cs.n_mem


@pytest.fixture(scope="module")
def million():
    import torch
    from crvhip import _lib as L, chain as CH, config as CF, detector as DT, synth_gpu as SG
    cfg = CF.ChainConfig(precision="fp32")
    cond, G = DT.nominal_conditions(), DT.sector_geom()
    S = SG.events("cosmic", N_EVENTS, seed=5)
    ds = SG.device_steps(S)
    ctx = CH.context(cfg, cond, G)
    ctx.set_params(cfg, 99, ds.evkey_hi, 32)
    base = ds.struct()
    whole = _run(torch, L, CH, ctx, base)
    # four calls over contiguous event ranges: step offsets from the sorted event column
    ev = S["event"].to(torch.int64)
    edges = torch.tensor([0, N_EVENTS // 4, N_EVENTS // 2, 3 * N_EVENTS // 4, N_EVENTS], device=ev.device)
    bounds = torch.searchsorted(ev, edges).cpu().tolist()
    parts = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        st = L.Steps(base.event + 4 * a, base.counter + 4 * a, base.edep + 8 * a, base.z + 8 * a, base.t + 8 * a, b - a)
        parts.append(_run(torch, L, CH, ctx, st))
    return torch, whole, parts


def test_chunked_calls_give_the_same_tables(million):
    torch, (pb, np_, cb, nc, nm), parts = million
    assert np_ > 40 * N_EVENTS and nc > 2 * N_EVENTS  # ~48 pulses and ~3 clusters per event
    assert sum(p[1] for p in parts) == np_ and sum(p[3] for p in parts) == nc and sum(p[4] for p in parts) == nm
    for k in PULSE_COLS:
        cat = torch.cat([p[0][k][:p[1]] for p in parts])
        assert torch.equal(cat, pb[k][:np_]), k
    for k in CLUSTER_COLS:
        m = 3 if k == "cl_pos" else 2 if k in ("cl_sidepe", "cl_sidetime", "cl_sidehits") else 1
        cat = torch.cat([p[2][k][:m * p[3]] for p in parts])
        assert torch.equal(cat, cb[k][:m * nc]), k
    # members (pulse indices within their event) equal; offsets shifted by the members of the earlier calls
    assert torch.equal(torch.cat([p[2]["mem_idx"][:p[4]] for p in parts]), cb["mem_idx"][:nm])
    off, m0 = [], 0
    for p in parts:
        off.append(p[2]["cl_mem_off"][:p[3]] + m0)
        m0 += p[4]
    assert torch.equal(torch.cat(off), cb["cl_mem_off"][:nc])
    assert int(cb["cl_mem_off"][nc]) == nm


def test_pulses_in_event_channel_time_order(million):
    torch, (pb, np_, cb, nc, nm), parts = million
    ev, ch, t = pb["ev"][:np_], pb["chan"][:np_].to(torch.int64), pb["time"][:np_]
    later_ev = ev[1:] > ev[:-1]
    same_ev = ev[1:] == ev[:-1]
    later_ch = ch[1:] > ch[:-1]
    same_ch = ch[1:] == ch[:-1]
    ok = later_ev | (same_ev & (later_ch | (same_ch & (t[1:] >= t[:-1]))))
    assert bool(ok.all())


def test_members_are_pulses_of_their_cluster_event(million):
    torch, (pb, np_, cb, nc, nm), parts = million
    off = cb["cl_mem_off"][:nc + 1]
    assert int(off[0]) == 0 and int(off[-1]) == nm and bool((off[1:] >= off[:-1]).all())
    # a member's row: the first pulse row of the cluster's event plus the member's index within the event
    ev = pb["ev"][:np_]
    cl_ev = cb["cl_event"][:nc]
    mem = cb["mem_idx"][:nm]
    cl_of_member = torch.repeat_interleave(torch.arange(nc, device=mem.device), off[1:] - off[:-1])
    first = torch.searchsorted(ev, cl_ev)  # the pulses are sorted by event
    assert torch.equal(ev[first], cl_ev)
    count = torch.searchsorted(ev, cl_ev, right=True) - first
    assert bool((mem >= 0).all()) and bool((mem < count[cl_of_member]).all())
    assert torch.equal(ev[first[cl_of_member] + mem], cl_ev[cl_of_member])
