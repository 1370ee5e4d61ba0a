"""The CRV response chain behind the reference's stage API (crvfastsim/chain.py), on the HIP kernels of libcrvhip.

    charges(steps, parts, geom, cond, cfg, rng, key=None, ...)   -> Q    reference chain.py:151
    digis(Q, cond, cfg, rng, key=None, ...)                     -> D    reference chain.py:234
    reco_waveforms(D, cond)                                     -> W    reference chain.py:273
    clusters(W, cond, geom, counters_cfg, cfg, ...)             -> ([out], timers)   reference chain.py:296

Tables are dicts of numpy arrays keyed as in the reference (labels = event key * CHAN_MULT + channel). Inputs are
uploaded to device buffers (torch tensors), the stage runs in HIP kernels, outputs come back as new host arrays;
the caller's arrays are never modified. Randomness is Philox keyed by an integer seed: `key` when given (the
reference's JAX key argument), otherwise one draw from `rng`. Errors follow the reference: ValueError for invalid
configuration, RuntimeError for incomplete device results (never truncated tables).

Steps: a DataFrame or dict with `event` (or `subrun` and `event`), `counter` (or `bar`), `edep`, and `z`/`t` (or the
reference's CrvStep columns z0, z1, t0, t1: the step midpoint along the counter and in time).
"""
import time as _time

import numpy as np

from . import _lib as L
from . import contract as CT
from . import detector as DT
from .config import CHAN_MULT, PERIOD, ChainConfig  # noqa: F401  (re-exported: the reference's chain names)
from .detector import Conditions, effective_status  # noqa: F401

_CTX = {}


def event_key(df):
    sub = np.asarray(df["subrun"], np.int64) if "subrun" in df else 0
    return (sub << 20) | np.asarray(df["event"], np.int64)


def context(cfg, cond, geom=None, device=0):
    """The cached device context for these configuration and conditions objects (geometry uploaded on demand)."""
    from . import template as TP
    k = (id(cfg), id(cond), device)
    ent = _CTX.get(k)
    if ent is None:
        _CTX.clear()
        ent = [L.Context(cfg, cond, geom, TP.load(cfg.waveform), device), cfg, cond, geom]
        _CTX[k] = ent                        # keep the keyed objects alive with the context
    elif geom is not None and ent[3] is not geom:
        ent[0].upload_geometry(geom, cfg.finder)
        ent[3] = geom
    return ent[0]


def _seed(key, rng):
    if key is not None:
        return int(key)
    if rng is None:
        return 0
    if isinstance(rng, (int, np.integer)):
        return int(rng)
    return int(rng.integers(2 ** 62))


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("crvhip needs a HIP device (MI355X); none is visible")
    return torch


def _dev(a, dtype):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda")


def trim():
    """Return the device workspace of every cached context (the tables of their last chain runs included); returns
    the bytes freed."""
    return sum(ent[0].trim() for ent in _CTX.values())


def _empty(n, dtype, may_trim=True):
    """A device buffer; when the device is full, the library's cached workspace is returned first and the
    allocation tried once more (not while a chain run's tables wait in the workspace to be fetched)."""
    torch = _torch()
    try:
        return torch.empty(max(int(n), 1), dtype=dtype, device="cuda")
    except torch.OutOfMemoryError:
        if not may_trim or not _CTX or trim() == 0:
            raise
        return torch.empty(max(int(n), 1), dtype=dtype, device="cuda")


def _host(t, n):
    return t[:n].cpu().numpy()


def step_table(steps, geom=None):
    """(evkey, counter, edep, z, t) of a step table, sorted stably by event key (returns also the order); rows with
    a non-finite time are left out. See step_columns for z and t."""
    ev = event_key(steps)
    counter, edep, z, t = step_columns(steps, geom)
    good = np.isfinite(t)
    order = np.flatnonzero(good)[np.argsort(ev[good], kind="stable")]
    return ev[order], counter[order], edep[order], z[order], t[order], order


def step_columns(steps, geom=None):
    """(counter, edep, z, t) per step row. z is the step's position along its counter from the counter centre: the `z`
    column, or, for the reference's CrvStep columns, the midpoint of (x0, y0, z0)-(x1, y1, z1) placed in its counter by
    a CRVGeometry `geom` (without one, z0 / z1 are taken as counter-local); t is the `t` column or the midpoint of
    t0 and t1."""
    counter = np.asarray(steps["counter"] if "counter" in steps else steps["bar"], np.int64)
    edep = np.asarray(steps["edep"], np.float64)
    if "z" in steps:
        z = np.asarray(steps["z"], np.float64)
    elif DT.is_crv_geometry(geom):
        if len(counter) and (counter.min() < 0 or counter.max() >= len(geom.sector)):
            raise ValueError(f"step bar {int(counter.max())} outside the geometry's {len(geom.sector)} bars")
        col = lambda e: np.column_stack([np.asarray(steps[f"{a}{e}"], float) for a in "xyz"])
        z = DT.step_longitudinal(geom, counter, col(0), col(1))
    else:
        z = 0.5 * (np.asarray(steps["z0"], float) + np.asarray(steps["z1"], float))
    t = np.asarray(steps["t"], np.float64) if "t" in steps else 0.5 * (np.asarray(steps["t0"], float) +
                                                                       np.asarray(steps["t1"], float))
    return counter, edep, z, t


def _check_response_geometry(geom, cfg):
    """The closed-form optics read every counter at +-cfg.optics.half_length: a CRVGeometry with other counter lengths
    is refused (a SectorGeom, or None, is the synthetic detector)."""
    if geom is None or isinstance(geom, DT.SectorGeom):
        return
    if not DT.is_crv_geometry(geom):
        raise ValueError(f"geom: expected a CRVGeometry, a SectorGeom or None, got {type(geom).__name__}")
    h = DT.bar_half_lengths(geom)
    bad = np.flatnonzero(np.abs(h - cfg.optics.half_length) > 1e-6 * cfg.optics.half_length)
    if len(bad):
        raise ValueError(f"geom: counter {int(bad[0])} has half length {h[bad[0]]:g} mm; the closed-form optics "
                         f"(cfg.optics.half_length) read counters of {cfg.optics.half_length:g} mm")


class DeviceSteps:
    """Steps resident on the device (SoA, 32 bytes per step), sorted by event."""

    def __init__(self, ev, counter, edep, z, t):
        torch = _torch()
        ev = np.asarray(ev, np.int64)
        if len(ev) and np.any(np.diff(ev) < 0):
            raise ValueError("steps must be sorted by event")
        hi = np.unique(ev >> 32)
        if len(hi) > 1:
            raise ValueError("steps of one call must share the event key's upper 32 bits")
        self.evkey_hi = int(hi[0]) if len(hi) else 0
        self.n = len(ev)
        self.event = torch.from_numpy((ev & 0xFFFFFFFF).astype(np.uint32).view(np.int32)).to("cuda")
        self.counter = _dev(np.asarray(counter).astype(np.int32), np.int32)
        self.edep, self.z, self.t = _dev(edep, np.float64), _dev(z, np.float64), _dev(t, np.float64)

    def struct(self):
        return L.Steps(L.ptr(self.event), L.ptr(self.counter), L.ptr(self.edep), L.ptr(self.z), L.ptr(self.t),
                       self.n)


def charges(steps, parts, geom, cond, cfg, rng, key=None, tabs=None, pmap=None, timers=None, pdg=None, caps=None):
    """CrvSteps -> SiPM charges: dict chan (label), time, charge, pe, photon (step row), sorted by (label, time).

    Without `tabs` the closed-form optics of cfg.optics apply: `pdg` must be None (`parts` is not read) and `geom` is
    None or a SectorGeom (synthetic detector), or a CRVGeometry whose counters all have the optics' length (steps
    placed by their world coordinates). With `tabs` ({name: crvhip.lookup.LookupTable}) the stock lookup-table optics
    apply (reference response/photons.py, the photons of crvhip.photons on the device): `geom` must be a CRVGeometry,
    the steps carry the CrvStep columns (bar, edep, t0, t1, x0..z1, px0, py0, pz0, p1, len) and the particle of each
    step comes from `pdg` or from `parts` (joined on the step's sp). `pmap` replaces cfg.photon_map for this call."""
    timers = timers if timers is not None else {}
    CT.validate_chain(cfg)
    if caps is not None:
        raise ValueError("caps: capacities need the batched chain")
    if tabs is not None:
        return _charges_lookup(steps, parts, geom, cond, cfg, rng, key, tabs, pmap, timers, pdg)
    if pdg is not None:
        raise ValueError("pdg: the closed-form optics have no particle-dependent (Cerenkov) yield; pass pdg=None")
    _check_response_geometry(geom, cfg)
    ev, counter, edep, z, t, order = step_table(steps, geom)
    if len(counter) and int(counter.max()) * 4 + 3 >= CHAN_MULT:
        raise ValueError(f"channel {int(counter.max()) * 4 + 3} >= CHAN_MULT {CHAN_MULT}")
    ctx = context(cfg, cond)
    ctx.use_photon_map(cfg, pmap)
    ds = DeviceSteps(ev, counter, edep, z, t)
    ctx.set_params(cfg, _seed(key, rng), ds.evkey_hi)
    t0 = _time.perf_counter()
    torch = _torch()
    cap = int(float(np.sum(edep)) * cfg.optics.pe_per_mev * 1.6 * 4 * 0.7) + 4096
    for _ in range(2):
        bufs = {k: _empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                                 ("charge", torch.float64), ("pe", torch.float64),
                                                 ("photon", torch.int64))}
        out = L.Charges(*[L.ptr(bufs[k]) for k in ("label", "time", "charge", "pe", "photon")], cap, 0, 0, 0)
        rc = ctx.lib.crv_charges_run(ctx.h, C_ref(ds.struct()), C_ref(out), None)
        if rc == L.CRV_E_CAPACITY:
            cap = out.n
            continue
        ctx.check(rc)
        break
    n = out.n
    Q = {"chan": _host(bufs["label"], n), "time": _host(bufs["time"], n), "charge": _host(bufs["charge"], n),
         "pe": _host(bufs["pe"], n), "photon": order[_host(bufs["photon"], n)] if n else np.zeros(0, np.int64)}
    timers["response_s"] = timers.get("response_s", 0.0) + _time.perf_counter() - t0
    timers["photons"] = timers.get("photons", 0) + int(out.n_photons)
    timers["charges"] = timers.get("charges", 0) + n
    if out.capped:
        timers["poisson_capped"] = timers.get("poisson_capped", 0) + int(out.capped)
    return Q


def _pdg_of(steps, parts, pdg):
    """pdg per step: given, or joined from the particle table on (run, subrun, event, sp) = (..., key), 0 if absent
    (reference chain.py:162-164)."""
    n = len(np.asarray(steps["bar"]))
    if pdg is not None:
        pdg = np.asarray(pdg)
        if pdg.shape != (n,) or (n and pdg.dtype.kind not in "iu"):
            raise ValueError("pdg must be an integer array aligned with the steps")
        return pdg.astype(np.int64)
    if parts is None:
        raise ValueError("the lookup optics need the particle of each step: pass pdg or parts")
    import pandas as pd
    keys = [k for k in ("run", "subrun", "event") if k in steps and k in parts]
    st = pd.DataFrame({k: np.asarray(steps[k]) for k in keys + ["sp"]})
    pt = pd.DataFrame({**{k: np.asarray(parts[k]) for k in keys}, "sp": np.asarray(parts["key"]),
                       "pdg": np.asarray(parts["pdg"])})
    return st.merge(pt, on=keys + ["sp"], how="left").pdg.fillna(0).to_numpy().astype(np.int64)


def _charges_lookup(steps, parts, geom, cond, cfg, rng, key, tabs, pmap, timers, pdg):
    """charges() through the lookup-table optics: crv_lookup_photons_run, then crv_sipm_run without dark noise or
    trap afterpulses (validate_chain refuses both, as the reference's contract does), photons after cfg.window[1]
    dropped (reference chain.py:193-213)."""
    from . import photons as PH
    if not isinstance(tabs, dict) or not tabs or not all(hasattr(T, "tables") and hasattr(T, "edges")
                                                         for T in tabs.values()):
        raise ValueError("tabs: expected a non-empty dict {name: crvhip.lookup.LookupTable}")
    if not DT.is_crv_geometry(geom):
        raise ValueError("geom: the lookup-table optics need a CRVGeometry (bar centres and axes)")
    pdg = _pdg_of(steps, parts, pdg)
    ev = event_key(steps)
    bar = np.asarray(steps["bar"], np.int64)
    nb = len(np.asarray(geom.sector))
    if len(bar) and (bar.min() < 0 or bar.max() >= nb):
        raise ValueError(f"step bar outside the geometry's {nb} bars")
    if nb * 4 > CHAN_MULT:
        raise ValueError(f"channel {nb * 4 - 1} >= CHAN_MULT {CHAN_MULT}: SiPM labels would overlap the next event")
    ctx = context(cfg, cond)
    ctx.use_photon_map(cfg, pmap)
    PH.upload(ctx, tabs, geom, cfg.response)
    ls = PH.LookupSteps(ev, steps, pdg)
    ctx.set_params(cfg, _seed(key, rng), ls.evkey_hi)
    t0 = _time.perf_counter()
    torch = _torch()
    empty = {k: np.zeros(0, np.int64 if k in ("chan", "photon") else np.float64)
             for k in ("chan", "time", "charge", "pe", "photon")}
    if ls.n == 0:
        return empty
    pb, po = PH.device_photons(ctx, ls)
    nph = po.n
    timers["photons_s"] = timers.get("photons_s", 0.0) + _time.perf_counter() - t0
    timers["photons"] = timers.get("photons", 0) + int(nph)
    if po.capped:
        timers["poisson_capped"] = timers.get("poisson_capped", 0) + int(po.capped)
    if nph == 0:
        return empty
    t0 = _time.perf_counter()
    s = cfg.sipm
    pl = L.PhotonList(L.ptr(pb["label"]), L.ptr(pb["time"]), L.ptr(pb["photon"]), L.ptr(pb["j"]), L.ptr(pb["rank"]),
                      nph)
    nz = L.SipmNoise(0.0, float(s.trap0_prob), float(s.trap1_prob), float(s.trap0_lifetime), float(s.trap1_lifetime),
                     0.0, float(cfg.window[1]), None, 0)
    cap = 2 * nph + 4096
    for _ in range(2):
        qb = {k: _empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                               ("charge", torch.float64), ("pe", torch.float64),
                                               ("photon", torch.int64))}
        qo = L.Charges(*[L.ptr(qb[k]) for k in ("label", "time", "charge", "pe", "photon")], cap, 0, 0, 0)
        rc = ctx.lib.crv_sipm_run(ctx.h, C_ref(pl), C_ref(nz), C_ref(qo), None)
        if rc == L.CRV_E_CAPACITY:
            cap = qo.n
            continue
        ctx.check(rc)
        break
    m = qo.n
    Q = {"chan": _host(qb["label"], m), "time": _host(qb["time"], m), "charge": _host(qb["charge"], m),
         "pe": _host(qb["pe"], m), "photon": ls.order[_host(qb["photon"], m)] if m else np.zeros(0, np.int64)}
    timers["sipm_s"] = timers.get("sipm_s", 0.0) + _time.perf_counter() - t0
    timers["charges"] = timers.get("charges", 0) + m
    return Q


def C_ref(x):
    import ctypes
    return ctypes.byref(x)


def digis(Q, cond, cfg, rng, key=None, timers=None, *, stock_scan=False, caps=None):
    """SiPM charges -> zero-suppressed digis: chan (label), start, tdc0, starttdc, adc [D, samples_zs], and the
    window records (record_adc [R, window_samples], record_label, record_tick)."""
    timers = timers if timers is not None else {}
    torch = _torch()
    lab = np.asarray(Q["chan"], np.int64)
    tm = np.asarray(Q["time"], np.float64)
    if len(lab) and (np.any(np.diff(lab) < 0) or np.any((np.diff(lab) == 0) & (np.diff(tm) < 0))):
        raise ValueError("charges must be grouped by label and time-ordered within it")
    ctx = context(cfg, cond)
    hi = np.unique((lab // CHAN_MULT) >> 32)
    ctx.set_params(cfg, _seed(key, rng), int(hi[0]) if len(hi) else 0)
    dl, dt, dq = _dev(lab, np.int64), _dev(tm, np.float64), _dev(Q["charge"], np.float64)
    t0 = _time.perf_counter()
    Sz, Wn = cfg.waveform.samples_zs, cfg.window_samples
    cap, rcap = max(64, len(lab)), max(64, len(lab))
    for _ in range(2):
        b = {"label": _empty(cap, torch.int64), "start": _empty(cap, torch.float64), "tdc0": _empty(cap, torch.float64),
             "starttdc": _empty(cap, torch.int64), "adc": _empty(cap * Sz, torch.int16),
             "record": _empty(cap, torch.int32), "sample": _empty(cap, torch.int32),
             "record_adc": _empty(rcap * Wn, torch.int16), "record_label": _empty(rcap, torch.int64),
             "record_tick": _empty(rcap, torch.int32)}
        out = L.Digis(L.ptr(b["label"]), L.ptr(b["start"]), L.ptr(b["tdc0"]), L.ptr(b["starttdc"]), L.ptr(b["adc"]),
                      L.ptr(b["record"]), L.ptr(b["sample"]), cap, 0, L.ptr(b["record_adc"]),
                      L.ptr(b["record_label"]), L.ptr(b["record_tick"]), rcap, 0, 0)
        rc = ctx.lib.crv_digis_run(ctx.h, L.ptr(dl), L.ptr(dt), L.ptr(dq), len(lab), C_ref(out), None)
        if rc == L.CRV_E_CAPACITY:
            cap, rcap = max(out.n, 1), max(out.n_records, 1)
            continue
        ctx.check(rc)
        break
    n, nr = out.n, out.n_records
    D = {"chan": _host(b["label"], n), "start": _host(b["start"], n), "tdc0": _host(b["tdc0"], n),
         "starttdc": _host(b["starttdc"], n), "adc": _host(b["adc"], n * Sz).reshape(n, Sz).astype(np.int16),
         "record": _host(b["record"], n), "sample": _host(b["sample"], n),
         "record_adc": _host(b["record_adc"], nr * Wn).view(np.uint16).reshape(nr, Wn),
         "record_label": _host(b["record_label"], nr), "record_tick": _host(b["record_tick"], nr),
         "overlong": int(out.overlong)}
    timers["waveforms_s"] = timers.get("waveforms_s", 0.0) + _time.perf_counter() - t0
    timers["digis"] = timers.get("digis", 0) + n
    return D


def reco_waveforms(D, cond):
    """Digis -> reco waveform rows (reference chain.py:273-286): ev, bar, sipm, chan, adc [W, L] (padded), n, start,
    ped, area, height, toff; sorted by (ev, chan, start)."""
    torch = _torch()
    lab = np.asarray(D["chan"], np.int64)
    st = np.asarray(D["starttdc"], np.int64)
    adc = np.asarray(D["adc"])
    o = np.lexsort((st, lab))
    lab, st, adc = lab[o], st[o], adc[o]
    Sz = adc.shape[1] if adc.ndim == 2 else 12
    nd = len(lab)
    if nd == 0:
        z = np.zeros(0, np.int64)
        return {"ev": z, "bar": z, "sipm": z, "chan": z, "adc": np.zeros((0, 0), np.int64), "n": z, "start": z,
                "ped": np.zeros(0), "area": np.zeros(0), "height": np.zeros(0), "toff": np.zeros(0)}
    ctx = _any_context()
    dl, ds_, da = _dev(lab, np.int64), _dev(st, np.int64), _dev(adc.astype(np.int16).ravel(), np.int16)
    b = {"ev": _empty(nd, torch.int64), "chan": _empty(nd, torch.int32), "start": _empty(nd, torch.int64),
         "nsamp": _empty(nd, torch.int32), "off": _empty(nd + 1, torch.int64), "adc": _empty(nd * Sz, torch.int16)}
    out = L.Waveforms(L.ptr(b["ev"]), L.ptr(b["chan"]), L.ptr(b["start"]), L.ptr(b["nsamp"]), L.ptr(b["off"]),
                      L.ptr(b["adc"]), nd, 0, nd * Sz, 0)
    ctx.check(ctx.lib.crv_reco_waveforms_run(ctx.h, L.ptr(dl), L.ptr(ds_), L.ptr(da), nd, C_ref(out), None))
    w = out.n
    ev, ch, start = _host(b["ev"], w), _host(b["chan"], w).astype(np.int64), _host(b["start"], w)
    n, off = _host(b["nsamp"], w).astype(np.int64), _host(b["off"], w + 1)
    flat = _host(b["adc"], out.n_samples).astype(np.int64)
    Lm = int(n.max())
    mat = np.zeros((w, Lm), np.int64)
    rows = np.repeat(np.arange(w), n)
    cols = np.arange(int(n.sum())) - np.repeat(off[:-1], n)
    mat[rows, cols] = flat
    return {"ev": ev, "bar": ch // 4, "sipm": ch % 4, "chan": ch, "adc": mat, "n": n, "start": start,
            "ped": np.asarray(cond.pedestal)[ch], "area": np.asarray(cond.pulsearea)[ch],
            "height": np.asarray(cond.pulseheight)[ch], "toff": np.asarray(cond.timeoffset)[ch]}


def _any_context():
    if not _CTX:
        raise RuntimeError("no device context: call charges/digis/clusters with a configuration first")
    return next(iter(_CTX.values()))[0]


PULSE_KEYS = ("wf", "seq", "pe", "peph", "time", "height", "beta", "chi2", "letime", "failed", "zerondf", "ped",
              "pedfromdb", "fit_A", "fit_mu", "fit_b")


def pulses(W, cond, cfg, geom=None):
    """Reco waveforms -> pulses (reference pulsefit.reconstruct_full + filter_channels + DB time offset)."""
    torch = _torch()
    ctx = context(cfg, cond, geom)
    ctx.set_params(cfg, 0)
    n = np.asarray(W["n"], np.int64)
    nw = len(n)
    if nw == 0:
        out = {k: np.zeros(0) for k in PULSE_KEYS}
        for k in ("wf", "seq"):
            out[k] = np.zeros(0, np.int64)
        for k in ("failed", "zerondf", "pedfromdb"):
            out[k] = np.zeros(0, bool)
        return out
    adc = np.asarray(W["adc"])
    mask = np.arange(adc.shape[1])[None, :] < n[:, None]
    flat = adc[mask].astype(np.int16)
    off = np.r_[0, np.cumsum(n)]
    chan = np.asarray(W["bar"]) * 4 + np.asarray(W["sipm"])
    toff = np.asarray(W["toff"] if "toff" in W else np.asarray(cond.timeoffset)[chan], np.float64)
    args = [_dev(np.asarray(W["ev"], np.int64), np.int64), _dev(chan.astype(np.int32), np.int32),
            _dev(np.asarray(W["start"], np.int64), np.int64), _dev(n.astype(np.int32), np.int32),
            _dev(off, np.int64), _dev(flat, np.int16)]
    cal = [_dev(np.asarray(W[k], np.float64), np.float64) for k in ("ped", "area", "height")] + [_dev(toff, np.float64)]
    cap = int(n.sum() // 2 + nw)
    b = {k: _empty(cap, torch.int64) for k in ("wf", "seq", "ev")}
    b["chan"] = _empty(cap, torch.int32)
    for k in L.PULSE_F64:
        b[k] = _empty(cap, torch.float64)
    for k in L.PULSE_U8:
        b[k] = _empty(cap, torch.uint8)
    out = L.Pulses(*[L.ptr(b[k]) for k in ("wf", "seq", "ev", "chan") + L.PULSE_F64 + L.PULSE_U8], cap, 0)
    ctx.check(ctx.lib.crv_pulses_run(ctx.h, *[L.ptr(a) for a in args], nw, *[L.ptr(a) for a in cal], C_ref(out), None))
    m = out.n
    P = {k: _host(b[k], m) for k in ("wf", "seq") + L.PULSE_F64}
    for k in L.PULSE_U8:
        P[k] = _host(b[k], m).astype(bool)
    P["fit_A"], P["fit_mu"], P["fit_b"] = P.pop("fit_a"), P.pop("fit_mu"), P.pop("fit_b")
    if "event_t0" in W and m:   # additive bunch correction, independent of DB calibration (chain.py:322-323)
        P["time"] = P["time"] + np.asarray(W["event_t0"], np.float64)[P["wf"]]
    return P


def finder_geometry(geom, counters_cfg):
    """The finder's SectorGeom: a SectorGeom as given, one built from a CRVGeometry and its parsed counters file
    (reference reco/coinc.py:82-110), or the synthetic detector's for None."""
    if geom is None:
        if counters_cfg is not None:
            raise ValueError("counters_cfg needs a CRVGeometry geom")
        return DT.sector_geom()
    if isinstance(geom, DT.SectorGeom):
        return geom
    if DT.is_crv_geometry(geom):
        if counters_cfg is None:
            raise ValueError("geom: a CRVGeometry needs its counters_cfg (the parsed counters file)")
        return DT.sector_geom_from(geom, counters_cfg)
    raise ValueError(f"geom: expected a CRVGeometry, a SectorGeom or None, got {type(geom).__name__}")


def clusters(W, cond, geom, counters_cfg, cfg, timers=None, detail=False, rng=None, caps=None):
    """Reco waveforms -> pulses -> coincidence clusters (reference chain.py:296-362 output, one dict in a list).
    With cfg.background, background pulses drawn from `rng` for every event of W join the signal pulses in (event,
    bar, SiPM, time) order before the finder (reference chain.py:337-342); `caps` (fixed device shapes of the
    reference's JAX path) is refused."""
    timers = timers if timers is not None else {}
    CT.validate_chain(cfg)
    if caps is not None:
        raise ValueError("caps: capacities need the batched chain")
    G = finder_geometry(geom, counters_cfg)
    t0 = _time.perf_counter()
    P = pulses(W, cond, cfg, G)
    timers["pulses_s"] = timers.get("pulses_s", 0.0) + _time.perf_counter() - t0
    wf = P["wf"]
    ev = np.asarray(W["ev"])[wf] if len(wf) else np.zeros(0, np.int64)
    o = np.lexsort((P["seq"], wf, ev))
    ev, wf = ev[o], wf[o]
    idx = np.arange(len(ev)) - np.searchsorted(ev, ev)
    bar, sipm = np.asarray(W["bar"])[wf], np.asarray(W["sipm"])[wf]
    time, pe, seq = P["time"][o], P["pe"][o], P["seq"][o]
    from . import background as BG
    smp = BG.sampler(cfg, rng if rng is not None else np.random.default_rng())
    if smp is not None:
        B = smp(np.unique(np.asarray(W["ev"])))
        ev, idx, bar, sipm, time, pe, seq, _ = BG.merge_pulses(ev, idx, bar, sipm, time, pe, seq, B)
        timers["background_pulses"] = timers.get("background_pulses", 0) + len(B["ev"])
    t0 = _time.perf_counter()
    out = find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G)
    timers["coinc_s"] = timers.get("coinc_s", 0.0) + _time.perf_counter() - t0
    timers["pulses"] = timers.get("pulses", 0) + len(ev)
    if detail:
        out["pulses"] = dict(ev=ev, idx=idx, bar=bar, sipm=sipm, seq=seq, pe=pe, time=time)
    else:
        for k in ("cl_pos", "mem_cl", "mem_idx", "mem_event", "cl_sidepe", "cl_sidetime", "cl_sidehits"):
            out.pop(k, None)
    return [out], timers


def find_clusters(ev, bar, sipm, time, pe, seq, cfg, cond, G):
    """Pulses sorted by (event, waveform, sequence) -> cluster table (reference chain.py:337-362 + coinc.find_event)."""
    torch = _torch()
    ctx = context(cfg, cond, G)
    ctx.set_params(cfg, 0)
    n = len(ev)
    names = ("cl_event", "cl_type", "cl_pe", "cl_start", "cl_end", "cl_slope", "cl_avgtime", "cl_poscalc")
    if n == 0:
        out = {k: np.zeros(0) for k in names}
        out.update(cl_pos=np.zeros((0, 3)), mem_cl=np.zeros(0, np.int64), mem_idx=np.zeros(0, np.int64),
                   mem_event=np.zeros(0, np.int64))
        return out
    args = [_dev(np.asarray(ev, np.int64), np.int64), _dev(np.asarray(bar, np.int32), np.int32),
            _dev(np.asarray(sipm, np.int32), np.int32), _dev(np.asarray(time, np.float64), np.float64),
            _dev(np.asarray(pe, np.float64), np.float64), _dev(np.asarray(seq, np.int64), np.int64)]
    b = {"cl_event": _empty(n, torch.int64), "cl_type": _empty(n, torch.int32), "cl_poscalc": _empty(n, torch.int32),
         "cl_sidehits": _empty(2 * n, torch.int32), "cl_mem_off": _empty(n + 1, torch.int64),
         "mem_idx": _empty(n, torch.int64)}
    for k in L.CLUSTER_F64:
        b[k] = _empty(3 * n if k == "cl_pos" else 2 * n if k in ("cl_sidepe", "cl_sidetime") else n, torch.float64)
    out = L.Clusters(L.ptr(b["cl_event"]), L.ptr(b["cl_type"]), L.ptr(b["cl_poscalc"]), L.ptr(b["cl_sidehits"]),
                     *[L.ptr(b[k]) for k in L.CLUSTER_F64], L.ptr(b["cl_mem_off"]), n, 0, L.ptr(b["mem_idx"]), n, 0)
    ctx.check(ctx.lib.crv_clusters_run(ctx.h, *[L.ptr(a) for a in args], n, C_ref(out), None))
    return cluster_table(b, out.n, out.n_mem)


def _pulse_bufs(n, may_trim=True):
    torch = _torch()
    b = {k: _empty(n, torch.int64, may_trim=may_trim) for k in ("wf", "seq", "ev")}
    b["chan"] = _empty(n, torch.int32, may_trim=may_trim)
    for k in L.PULSE_F64:
        b[k] = _empty(n, torch.float64, may_trim=may_trim)
    for k in L.PULSE_U8:
        b[k] = _empty(n, torch.uint8, may_trim=may_trim)
    s = L.Pulses(*[L.ptr(b[k]) for k in ("wf", "seq", "ev", "chan") + L.PULSE_F64 + L.PULSE_U8], n, 0)
    return b, s


def _cluster_bufs(nc, nm, may_trim=True):
    torch = _torch()
    e = lambda n, dt: _empty(n, dt, may_trim=may_trim)
    b = {"cl_event": e(nc, torch.int64), "cl_type": e(nc, torch.int32), "cl_poscalc": e(nc, torch.int32),
         "cl_sidehits": e(2 * nc, torch.int32), "cl_mem_off": e(nc + 1, torch.int64), "mem_idx": e(nm, torch.int64)}
    for k in L.CLUSTER_F64:
        b[k] = e(3 * nc if k == "cl_pos" else 2 * nc if k in ("cl_sidepe", "cl_sidetime") else nc, torch.float64)
    s = L.Clusters(L.ptr(b["cl_event"]), L.ptr(b["cl_type"]), L.ptr(b["cl_poscalc"]), L.ptr(b["cl_sidehits"]),
                   *[L.ptr(b[k]) for k in L.CLUSTER_F64], L.ptr(b["cl_mem_off"]), nc, 0, L.ptr(b["mem_idx"]), nm, 0)
    return b, s


def _fetch(addr, n, dtype):
    """n elements at a device address of the library's workspace, copied to a new host array (hipMemcpy of the HIP
    runtime the process already runs; the soname resolves to the loaded copy)."""
    import ctypes
    out = np.empty(max(int(n), 0), dtype)
    if n:
        hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rc = hip.hipMemcpy(out.ctypes.data, addr, out.nbytes, 2)  # hipMemcpyDeviceToHost
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


def run_chain(steps, cond, cfg, key=0, geom=None, batch=None, timers=None, records=False):
    """CrvSteps -> pulses and coincidence clusters in one device call (the reference drives the same stages one by one,
    analyses/runchain.py:517-530 -> chain.py:151-362).

    cfg.precision "fp64": the plain FP64 chain (one launch per stage, equal to the oracle bit for bit); "fp32": the
    batched FP32 chain (three fused kernels, FP64 only for absolute times and the fit's 3 x 3 solve), equal to the
    FP64 chain statistically. `batch` holds crv_batch_opts fields for the FP32 chain (grid, tile, pulse_slots,
    sort_tasks, log_budget, pipeline). Returns {"pulses": {...}, "clusters": cluster table, "timing": {...}} with pulses in (event, channel,
    time) order and `wf` the pulse's reco-waveform number; with records=True (FP32 chain) also "records", the ADC
    records of the clusters with digis sorted by (label, FEB tick)."""
    import ctypes
    CT.validate_chain(cfg)
    if cfg.background is not None:
        raise ValueError("background: the one-call chain has no background pulses; run the stages (clusters) instead")
    G = geom if isinstance(geom, DT.SectorGeom) else finder_geometry(None, None) if geom is None else None
    if G is None:
        raise ValueError("geom: run_chain takes a SectorGeom or None (build one with finder_geometry)")
    ev, counter, edep, z, t, _ = step_table(steps)
    ds = DeviceSteps(ev, counter, edep, z, t)
    ctx = context(cfg, cond, G)
    ctx.set_params(cfg, _seed(key, None), ds.evkey_hi, 32 if cfg.precision == "fp32" else 64)
    st = ds.struct()
    tm = L.Timing()
    guess = max(1024, 4 * ds.n)
    pb, ps = _pulse_bufs(guess)
    cb, cs = _cluster_bufs(guess, guess)
    opts = L.BatchOpts()
    for k, v in (batch or {}).items():
        setattr(opts, k, v)
    if cfg.precision == "fp32":
        rc = ctx.lib.crv_chain_run_batched(ctx.h, ctypes.byref(st), ctypes.byref(opts), ctypes.byref(ps),
                                           ctypes.byref(cs), ctypes.byref(tm), None)
    elif cfg.precision == "fp64":
        rc = ctx.lib.crv_chain_run(ctx.h, ctypes.byref(st), ctypes.byref(ps), ctypes.byref(cs), ctypes.byref(tm), None)
    else:
        raise ValueError(f"precision must be fp64 or fp32, not {cfg.precision!r}")
    if rc == -3:  # CRV_E_CAPACITY: the sizes are in n; fetch again into larger buffers
        pb, ps = _pulse_bufs(ps.n, may_trim=False)
        cb, cs = _cluster_bufs(cs.n, cs.n_mem, may_trim=False)
        rc = ctx.lib.crv_chain_fetch(ctx.h, ctypes.byref(ps), ctypes.byref(cs), None)
    ctx.check(rc)
    out = _tables(pb, ps, cb, cs, tm, opts if cfg.precision == "fp32" else None, timers)
    if records and cfg.precision == "fp32":  # the records of the clusters with digis, by (label, tick)
        nr, Wn = opts.n_records, cfg.window_samples
        lab, tick = _fetch(opts.record_label, nr, np.int64), _fetch(opts.record_tick, nr, np.int32)
        adc = _fetch(opts.record_adc, nr * Wn, np.uint16).reshape(nr, Wn)
        keep = np.flatnonzero(lab >= 0)
        o = keep[np.lexsort((tick[keep], lab[keep]))]
        out["records"] = {"label": lab[o], "tick": tick[o], "adc": adc[o]}
    return out


def _tables(pb, ps, cb, cs, tm, opts=None, timers=None):
    n = ps.n
    P = {k: _host(pb[k], n) for k in ("wf", "seq", "ev", "chan") + L.PULSE_F64 + L.PULSE_U8}
    for k in L.PULSE_U8:
        P[k] = P[k].astype(bool)
    P["bar"], P["sipm"] = P["chan"] // 4, P["chan"] % 4
    timing = {f: getattr(tm, f) for f, _ in L.Timing._fields_}
    if opts is not None:
        timing.update(n_records=opts.n_records, n_waveforms=opts.n_waveforms, overlong=opts.overlong,
                      capped=opts.capped, pulse_slots_used=opts.pulse_slots_used, log_passes=opts.log_passes)
    if timers is not None:
        timers.update(timing)
    return {"pulses": P, "clusters": cluster_table(cb, cs.n, cs.n_mem), "timing": timing}


class ChainGraph:
    """The FP32 batched chain captured into one HIP graph (crv_graph_*) for calls of up to `max_steps` steps, sized on
    the `sample` steps. `run(steps)` replays it (calls beyond its capacities run uncaptured, same results)."""

    def __init__(self, sample, cond, cfg, key=0, geom=None, batch=None, max_steps=None, margin=0.25):
        import dataclasses
        CT.validate_chain(cfg)
        if cfg.background is not None:
            raise ValueError("background: the captured chain has no background pulses; run the stages instead")
        cfg = cfg if cfg.precision == "fp32" else dataclasses.replace(cfg, precision="fp32")
        self.cfg, self.cond = cfg, cond
        self.G = geom if geom is not None else DT.sector_geom()
        ev, counter, edep, z, t, _ = step_table(sample)
        ds = DeviceSteps(ev, counter, edep, z, t)
        self.ctx = context(cfg, cond, self.G)
        self.key = _seed(key, None)
        self.evkey_hi = ds.evkey_hi
        self.ctx.set_params(cfg, self.key, ds.evkey_hi, 32)
        opts = L.BatchOpts()
        for k, v in {"tile": 64, "sort_tasks": 1, **(batch or {})}.items():
            setattr(opts, k, v)
        self.graph = L.Graph(self.ctx, ds.struct(), max_steps or max(ds.n, 1), opts, margin)

    def run(self, steps, timers=None):
        import ctypes
        ev, counter, edep, z, t, _ = step_table(steps)
        ds = DeviceSteps(ev, counter, edep, z, t)
        if ds.evkey_hi != self.evkey_hi:
            raise ValueError("steps: the graph was captured for another upper half of the event key")
        self.ctx.set_params(self.cfg, self.key, self.evkey_hi, 32)   # the captured parameters
        tm = L.Timing()
        guess = max(1024, 4 * ds.n)
        pb, ps = _pulse_bufs(guess)
        cb, cs = _cluster_bufs(guess, guess)
        rc = self.graph.launch(ds.struct(), ps, cs, tm)
        if rc == L.CRV_E_CAPACITY:
            pb, ps = _pulse_bufs(ps.n, may_trim=False)
            cb, cs = _cluster_bufs(cs.n, cs.n_mem, may_trim=False)
            self.graph.fetch(ps, cs)
        out = _tables(pb, ps, cb, cs, tm, None, timers)
        out["graph"] = self.graph.stats()
        return out

    def close(self):
        self.graph.close()


def cluster_table(b, nc, nm):
    mo = _host(b["cl_mem_off"], nc + 1)
    cnt = np.diff(mo) if nc else np.zeros(0, np.int64)
    ev = _host(b["cl_event"], nc)
    out = {"cl_event": ev, "cl_type": _host(b["cl_type"], nc).astype(np.int64), "cl_pe": _host(b["cl_pe"], nc),
           "cl_start": _host(b["cl_start"], nc), "cl_end": _host(b["cl_end"], nc),
           "cl_slope": _host(b["cl_slope"], nc), "cl_avgtime": _host(b["cl_avgtime"], nc),
           "cl_poscalc": _host(b["cl_poscalc"], nc).astype(bool), "cl_pos": _host(b["cl_pos"], 3 * nc).reshape(nc, 3),
           "cl_sidepe": _host(b["cl_sidepe"], 2 * nc).reshape(nc, 2),
           "cl_sidetime": _host(b["cl_sidetime"], 2 * nc).reshape(nc, 2),
           "cl_sidehits": _host(b["cl_sidehits"], 2 * nc).reshape(nc, 2).astype(np.int64),
           "mem_cl": np.repeat(np.arange(nc), cnt), "mem_idx": _host(b["mem_idx"], nm),
           "mem_event": np.repeat(ev, cnt)}
    return out
