"""Synthetic CrvStep tables generated on the device (torch), for the large benchmark configurations.

Same distributions as crvhip.synth (cosmic tracks, the uneven 90/9/1 mixture, the Poisson background overlay) drawn
with torch's device generator instead of numpy's PCG64, so that tens of millions of events are generated in seconds
and stay resident in HBM. The tables are sorted by event and returned as device tensors in the layout the C ABI takes
(event u32 as int32 bits, counter i32, edep / z / t f64). Input generation is never timed.
"""
import numpy as np

from . import detector as DT
from .synth import LAYER_DT, T_RANGE, landau


def _gen(seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def _rep(torch, values, counts):
    return torch.repeat_interleave(values, counts)


def _landau(torch, g, n, device):
    L = landau()
    c = torch.as_tensor(L.c, dtype=torch.float64, device=device)
    x = torch.as_tensor(L.x, dtype=torch.float64, device=device)
    u = c[0] + (c[-1] - c[0]) * torch.rand(n, generator=g, dtype=torch.float64, device=device)
    i = torch.searchsorted(c, u).clamp(1, len(c) - 1)
    w = (u - c[i - 1]) / (c[i] - c[i - 1])
    return L.mpv + L.width * (x[i - 1] + w * (x[i] - x[i - 1]) - L.mode)


def _uniform(torch, g, n, lo, hi, device):
    return lo + (hi - lo) * torch.rand(n, generator=g, dtype=torch.float64, device=device)


def _randint(torch, g, n, hi, device):
    return torch.randint(0, int(hi), (n,), generator=g, device=device)


def _tracks(torch, g, ev_tr, module, col, z, t, device):
    T = len(ev_tr)
    layer = torch.arange(DT.N_LAYERS, device=device).repeat(T)
    module4, col4 = module.repeat_interleave(4), col.repeat_interleave(4)
    return dict(event=ev_tr.repeat_interleave(4), counter=(module4 * DT.N_LAYERS + layer) * DT.N_COLS + col4,
                edep=_landau(torch, g, 4 * T, device), z=z.repeat_interleave(4),
                t=t.repeat_interleave(4) + LAYER_DT * layer.double())


def _cosmic(torch, g, events, device):
    ntr = torch.poisson(torch.full((len(events),), 3.0, device=device), generator=g).long()
    ev_tr = _rep(torch, events, ntr)
    T = len(ev_tr)
    return _tracks(torch, g, ev_tr, _randint(torch, g, T, DT.N_MODULES, device), _randint(torch, g, T, DT.N_COLS, device),
                   _uniform(torch, g, T, -DT.HALF_LENGTH, DT.HALF_LENGTH, device),
                   _uniform(torch, g, T, 0.0, T_RANGE, device), device)


def _showers(torch, g, events, device):
    n = len(events)
    ntr = torch.poisson(torch.full((n,), 40.0, device=device), generator=g).long()
    m0 = _randint(torch, g, n, DT.N_MODULES - 7, device)
    t0 = _uniform(torch, g, n, 0.0, T_RANGE, device)
    ev_tr = _rep(torch, events, ntr)
    T = len(ev_tr)
    return _tracks(torch, g, ev_tr, _rep(torch, m0, ntr) + _randint(torch, g, T, 8, device),
                   _randint(torch, g, T, DT.N_COLS, device), _uniform(torch, g, T, -DT.HALF_LENGTH, DT.HALF_LENGTH, device),
                   _rep(torch, t0, ntr) + 0.5 * torch.randn(T, generator=g, dtype=torch.float64, device=device), device)


def _dense(torch, g, events, device):
    n = len(events)
    ns = torch.floor(torch.exp(_uniform(torch, g, n, np.log(200.0), np.log(2001.0), device))).long()
    m0 = _randint(torch, g, n, DT.N_MODULES - 1, device)
    z0 = _uniform(torch, g, n, -DT.HALF_LENGTH, DT.HALF_LENGTH, device)
    t0 = _uniform(torch, g, n, 0.0, T_RANGE, device)
    S = int(ns.sum())
    layer = _randint(torch, g, S, DT.N_LAYERS, device)
    module = _rep(torch, m0, ns) + _randint(torch, g, S, 2, device)
    z = (_rep(torch, z0, ns) + 300.0 * torch.randn(S, generator=g, dtype=torch.float64, device=device))
    return dict(event=_rep(torch, events, ns),
                counter=(module * DT.N_LAYERS + layer) * DT.N_COLS + _randint(torch, g, S, DT.N_COLS, device),
                edep=-4.0 * torch.log1p(-torch.rand(S, generator=g, dtype=torch.float64, device=device)),
                z=z.clamp(-DT.HALF_LENGTH, DT.HALF_LENGTH), t=_rep(torch, t0, ns) + LAYER_DT * layer.double())


def _background(torch, g, events, rate, device):
    n = len(events)
    nb = torch.poisson(torch.full((n,), rate * DT.N_COUNTERS, device=device), generator=g).long()
    B = int(nb.sum())
    return dict(event=_rep(torch, events, nb), counter=_randint(torch, g, B, DT.N_COUNTERS, device),
                edep=-0.8 * torch.log1p(-torch.rand(B, generator=g, dtype=torch.float64, device=device)),
                z=_uniform(torch, g, B, -DT.HALF_LENGTH, DT.HALF_LENGTH, device),
                t=_uniform(torch, g, B, 0.0, T_RANGE, device))


def _block(torch, g, workload, ev, rate, device):
    parts = []
    if workload == "cosmic":
        parts.append(_cosmic(torch, g, ev, device))
    elif workload in ("uneven", "overlay"):
        kind = torch.multinomial(torch.tensor([0.90, 0.09, 0.01], device=device), len(ev), replacement=True,
                                 generator=g)
        parts += [_cosmic(torch, g, ev[kind == 0], device), _showers(torch, g, ev[kind == 1], device),
                  _dense(torch, g, ev[kind == 2], device)]
        if workload == "overlay":
            parts.append(_background(torch, g, ev, rate, device))
    else:
        raise ValueError(f"unknown workload {workload}")
    parts = [p for p in parts if len(p["event"])]
    S = {k: torch.cat([p[k] for p in parts]) for k in ("event", "counter", "edep", "z", "t")}
    o = torch.sort(S["event"], stable=True).indices
    return {"event": S["event"][o].to(torch.int32), "counter": S["counter"][o].to(torch.int32),
            "edep": S["edep"][o], "z": S["z"][o], "t": S["t"][o]}


def events(workload, n_events, seed=0, device="cuda", rate=0.5, block=None):
    """Steps of n_events events of `workload` (cosmic, uneven or overlay) on the device, sorted by event, generated
    in blocks of events (bounded temporaries): dict(event i32, counter i32, edep, z, t f64) of torch tensors."""
    import torch
    g = _gen(seed, device)
    block = block or (20_000 if workload == "overlay" else 500_000)
    blocks = [_block(torch, g, workload, torch.arange(a, min(a + block, n_events), device=device), rate, device)
              for a in range(0, n_events, block)]
    if not blocks:
        blocks = [{k: torch.zeros(0, dtype=torch.int32 if k in ("event", "counter") else torch.float64, device=device)
                   for k in ("event", "counter", "edep", "z", "t")}]
    S = {}
    for k in ("event", "counter", "edep", "z", "t"):
        S[k] = torch.cat([b_[k] for b_ in blocks])
        for b_ in blocks:
            del b_[k]
    S["n_events"] = n_events
    return S


def device_steps(S):
    """DeviceSteps over the device tables (event as u32 bits in int32, counter int32); events < 2^31."""
    from .chain import DeviceSteps
    ds = DeviceSteps.__new__(DeviceSteps)
    ds.evkey_hi = 0
    ds.n = len(S["event"])
    ds.event, ds.counter = S["event"], S["counter"]
    ds.edep, ds.z, ds.t = S["edep"], S["z"], S["t"]
    return ds


def host_sample(S, n_events):
    """The first n_events events as host numpy arrays (event key = event)."""
    import torch
    cut = int(torch.searchsorted(S["event"], torch.tensor([n_events], dtype=torch.int32, device=S["event"].device)).item())
    out = {k: S[k][:cut].cpu().numpy() for k in ("event", "counter", "edep", "z", "t")}
    out["event"] = out["event"].astype("int64")
    out["counter"] = out["counter"].astype("int64")
    out["evkey"] = out["event"].copy()
    return out
