"""CRV photons at the SiPMs from CrvSteps through the stock v6 lookup tables, on the device.

Mirrors the reference's response/photons.py interface: SectorConfig / EXTRACTED_V05 (:31-43, in crvhip.config),
particle_props (:65-77), beta_of (:80-85), cerenkov_per_mm (:88-92), find_bin (:95-99), table_propagation
(:230-254), bar_keys / bar_ratios (:257-289), step_inputs (:292-302) and make_photons (:359-371). The per-step
host arithmetic stays in numpy where the reference's does (particle lookup, momentum magnitudes, per-bar ratios,
the per-table speed and attenuation fits); the substep loop, its Binomial arrivals and delay draws and the photon
numbers run in libcrvhip (crv_lookup_photons_run, FP64, draws keyed by Philox on the photon's identity).
"""
import numpy as np

from . import _lib as L
from .config import EXTRACTED_V05, SectorConfig  # noqa: F401  (reference names)

PRECISION_MM = 0.1

# pdg -> (mass MeV, charge e): the particles a CrvStep carries (PDG 2024 masses); ions 100ZZZAAAI in particle_props
PDG = {11: (0.51099895, -1), 13: (105.6583755, -1), 22: (0.0, 0), 211: (139.57039, 1), 111: (134.9768, 0),
       321: (493.677, 1), 130: (497.611, 0), 310: (497.611, 0), 2212: (938.272088, 1), 2112: (939.565420, 0),
       12: (0.0, 0), 14: (0.0, 0), 3122: (1115.683, 0), 3222: (1189.37, 1), 3112: (1197.449, -1)}
SELF_CONJUGATE = (22, 111, 130, 310)
AMU = 931.49410242


def particle_props(pdg):
    """Mass [MeV], charge [e] and known flag per pdg code; antiparticles flip the charge; ions 100ZZZAAAI have
    charge Z and mass A * amu; unknown codes give (0, 0, False)."""
    pdg = np.asarray(pdg, np.int64)
    m, q = np.zeros(len(pdg)), np.zeros(len(pdg))
    known = np.zeros(len(pdg), bool)
    for code, (mm, qq) in PDG.items():
        for c, sign in ((code, 1), (-code, -1)):
            if c == -code and code in SELF_CONJUGATE:
                continue
            s = pdg == c
            m[s], q[s], known[s] = mm, sign * qq, True
    ion = np.abs(pdg) >= 1_000_000_000
    a = (np.abs(pdg[ion]) // 10) % 1000
    z = (np.abs(pdg[ion]) // 10000) % 1000
    m[ion], q[ion], known[ion] = a * AMU, z * np.sign(pdg[ion]), True
    return m, q, known


def beta_of(p0, p1, mass):
    """beta from the mean of the start and end energies; massless -> 1."""
    p0, p1, mass = (np.asarray(a, np.float64) for a in (p0, p1, mass))
    e = 0.5 * (np.sqrt(p0 * p0 + mass * mass) + np.sqrt(p1 * p1 + mass * mass))
    g = e / np.where(mass > 0, mass, 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(mass > 0, np.sqrt(np.clip(1.0 - 1.0 / (g * g), 0.0, 1.0)), 1.0)


def cerenkov_per_mm(beta, charge, cmap):
    """Cerenkov photons per mm: linear in beta from (0, 0) through the map's points, flat beyond its end, x |q|."""
    b = np.r_[0.0, np.asarray(cmap, np.float64)[:, 0]]
    n = np.r_[0.0, np.asarray(cmap, np.float64)[:, 1]]
    return np.interp(beta, b, n, right=n[-1]) * np.abs(charge)


def find_bin(edges, v):
    """The first i with edges[i] <= v <= edges[i+1] (both edges inclusive); -1 outside."""
    edges = np.asarray(edges, np.float64)
    i = np.searchsorted(edges, v, side="left")
    i = np.where(np.asarray(v) == edges[0], 1, i)
    return np.where((i >= 1) & (i <= len(edges) - 1), i - 1, -1)


_PROP = {}


def table_propagation(T):
    """(inverse signal speed [ns/mm], attenuation length [mm]) of a table: the slope of the 10 % quantile of its
    scintillation delays against the distance d to the SiPM end, and -1 / the slope of the log of its summed arrival
    probability against d, over 0.2 < d / half length < 1.8 (Offline MakeCrvPhotons::FindTablePropagation)."""
    key = (id(T), T.name, T.half_length)
    if key in _PROP and _PROP[key][0] is T:
        return _PROP[key][1]
    E, S = T.edges, T.tables["scint"]
    nx, ny, nz = len(E["x"]) - 1, len(E["y"]) - 1, len(E["z"]) - 1
    p = S.prob.astype(np.float64).reshape(nx, ny, nz)
    tot_w = S.delay_sum.astype(np.float64).reshape(nx, ny, nz)
    w = np.where(tot_w > 0, p / np.where(tot_w > 0, tot_w, 1.0), 0.0)
    h = np.einsum("xyz,xyzj->zj", w, S.delays.astype(np.float64).reshape(nx, ny, nz, -1))
    d = 0.5 * (E["z"][1:] + E["z"][:-1]) + T.half_length
    tot = h.sum(1)
    ok = tot > 0
    q10 = np.argmax(np.cumsum(h, 1) >= 0.1 * tot[:, None], axis=1)
    inv_v = np.polyfit(d[ok], q10[ok].astype(float), 1)[0]
    P = p.sum((0, 1))
    a = (P > 0) & (d > 0.2 * T.half_length) & (d < 1.8 * T.half_length)
    sl = np.polyfit(d[a], np.log(P[a]), 1)[0]
    if not sl < 0:
        raise ValueError(f"cannot determine the attenuation length of {T.name}")
    out = (float(inv_v), float(-1.0 / sl))
    _PROP[key] = (T, out)
    return out


def bar_keys(geom):
    """Per bar: (sector key, module key); the module key is the sector name + module index when the sector has
    several modules."""
    names = np.array(geom.sector_names, dtype=object)
    sector = np.asarray(geom.sector, np.int64)
    module = np.asarray(geom.module, np.int64)
    nmod = np.zeros(len(names), np.int64)
    np.maximum.at(nmod, sector, module + 1)
    sk = names[sector]
    mk = np.where(nmod[sector] > 1, sk + module.astype(str).astype(object), sk)
    return sk, mk


def bar_ratios(values, geom):
    """Per-bar ratios from {key: ratio}: key TAG (every sector whose name starts with M), a sector name or a module
    key, the more specific winning; unlisted bars 1."""
    nbars = len(np.asarray(geom.sector))
    out = np.ones(nbars)
    if not values:
        return out
    sk, mk = bar_keys(geom)
    sectors, modules = set(sk), set(mk)
    bad = sorted(set(values) - sectors - modules - {"TAG"})
    if bad:
        raise KeyError(f"unknown propagation keys {bad}; known: {sorted(sectors | modules | {'TAG'})}")
    if "TAG" in values:
        out[np.array([str(n).startswith("M") for n in sk], bool)] = values["TAG"]
    for k, v in values.items():
        if k in sectors:
            out[sk == k] = v
    for k, v in values.items():
        if k in modules and k not in sectors:
            out[mk == k] = v
    if not np.all(out > 0):
        raise ValueError("propagation ratios must be positive")
    return out


def _col(steps, k):
    v = steps[k]
    return np.asarray(v.to_numpy() if hasattr(v, "to_numpy") else v)


def step_inputs(steps, pdg, geom):
    """Per step: local start and end (thickness, width, length) in its bar, beta, charge and the known-pdg flag."""
    bar = _col(steps, "bar").astype(np.int64)
    l2w = np.asarray(geom.local_to_world, np.int64)[bar]
    c = np.asarray(geom.center, np.float64)[bar]
    r = np.arange(len(bar))[:, None]
    s = np.column_stack([_col(steps, f"{a}0").astype(float) for a in "xyz"]) - c
    e = np.column_stack([_col(steps, f"{a}1").astype(float) for a in "xyz"]) - c
    mass, charge, known = particle_props(pdg)
    p0 = np.sqrt(_col(steps, "px0") ** 2 + _col(steps, "py0") ** 2 + _col(steps, "pz0") ** 2)
    return s[r, l2w], e[r, l2w], beta_of(p0, _col(steps, "p1"), mass), charge, known


def _tables_of(tables):
    """The tables as a list and a name -> index map; names with and without the LookupTable_ prefix."""
    names = sorted(tables)
    idx = {}
    for i, n in enumerate(names):
        idx[n] = i
        idx[n.replace("LookupTable_", "")] = i
    return [tables[n] for n in names], idx


def bar_arrays(geom, rcfg, tables):
    """Per-bar inputs of the device substep loop: table index (-1: a sector the configuration does not simulate),
    reflector, yield, speed and attenuation ratios, centre, local_to_world. Raises like the reference's iter_rows
    for a sector without configuration, a missing table or a reflector that does not match its table."""
    tl, idx = _tables_of(tables)
    names = list(geom.sector_names)
    sector = np.asarray(geom.sector, np.int64)
    nb = len(sector)
    tab = np.full(nb, -1, np.int32)
    refl = np.zeros(nb, np.int32)
    yld = np.zeros(nb)
    present = set(np.unique(sector).tolist())
    for si, n in enumerate(names):
        if si not in present:
            continue
        if n not in rcfg.sectors:
            raise ValueError(f"sector {n!r} has no entry in ResponseConfig.sectors")
        sc = rcfg.sectors[n]
        if sc.table not in idx:
            raise ValueError(f"sector {n!r}: lookup table {sc.table!r} not among the given tables {sorted(tables)}")
        T = tl[idx[sc.table]]
        if abs(int(sc.reflector)) != int(T.reflector):
            raise ValueError(f"reflector {sc.reflector} does not match lookup table {sc.table} ({T.reflector})")
        m = sector == si
        tab[m], refl[m], yld[m] = idx[sc.table], int(sc.reflector), float(sc.yield_per_mev)
    return {"table": tab, "reflector": refl, "yield": yld,
            "speed_ratio": bar_ratios(rcfg.speed_ratio, geom).astype(np.float64),
            "atten_ratio": bar_ratios(rcfg.atten_ratio, geom).astype(np.float64),
            "center": np.ascontiguousarray(geom.center, np.float64),
            "l2w": np.ascontiguousarray(geom.local_to_world, np.int32)}, tl


def upload(ctx, tables, geom, rcfg):
    """Upload the lookup tables and the per-bar arrays to a context (kept until the next upload); returns the bar
    arrays and the (inv_v, att) of each table in upload order."""
    bars, tl = bar_arrays(geom, rcfg, tables)
    for T in tl:
        if T.fiber_decay_time > 0:
            raise ValueError(f"{T.name}: fibre decay time {T.fiber_decay_time} > 0 (fibre-emission decays) is not "
                             "supported; the v6 tables carry 0")
    props = [table_propagation(T) for T in tl]
    keep = []

    def h(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    ct = (L.LookupTableC * len(tl))()
    for i, (T, (inv_v, att)) in enumerate(zip(tl, props)):
        c = ct[i]
        for k in ("half_thickness", "half_width", "half_length", "fiber_separation", "hole_radius_x", "hole_radius_y",
                  "fiber_radius", "corner_radius"):
            setattr(c, k, float(getattr(T, k)))
        c.inv_speed, c.atten_length = inv_v, att
        for k, name in enumerate(("x", "y", "z", "beta", "theta", "phi", "r")):
            c.edges[k] = h(T.edges[name], np.float64)
            c.n_edges[k] = len(T.edges[name])
        for k, m in enumerate((T.ck_scint, T.ck_fiber)):
            m = np.asarray(m, np.float64).reshape(-1, 2)
            c.ck[k] = h(m, np.float64)
            c.n_ck[k] = len(m)
        for k, name in enumerate(("scint", "ckscint", "ckfiber")):
            S = T.tables[name]
            c.prob[k] = h(S.prob, np.float32)
            c.delays[k] = h(S.delays, np.uint8)
            c.dmax[k] = S.delays.shape[1]
    cb = L.LookupBars(len(bars["table"]), h(bars["table"], np.int32), h(bars["reflector"], np.int32),
                      h(bars["yield"], np.float64), h(bars["speed_ratio"], np.float64),
                      h(bars["atten_ratio"], np.float64), h(bars["center"], np.float64), h(bars["l2w"], np.int32))
    import ctypes
    ctx.check(ctx.lib.crv_lookup_upload(ctx.h, ct, len(tl), ctypes.byref(cb)))
    return bars, props


class LookupSteps:
    """Steps for the lookup optics, resident on the device and sorted by event key (the sort order `order`)."""

    def __init__(self, evkey, steps, pdg):
        from .chain import _dev
        import torch
        ev = np.asarray(evkey, np.int64)
        t0, t1 = _col(steps, "t0").astype(np.float64), _col(steps, "t1").astype(np.float64)
        self.order = np.argsort(ev, kind="stable")
        o = self.order
        ev = ev[o]
        hi = np.unique(ev >> 32)
        if len(hi) > 1:
            raise ValueError("steps of one call must share the event key's upper 32 bits")
        self.evkey_hi = int(hi[0]) if len(hi) else 0
        self.evkey = ev
        self.n = len(ev)
        mass, charge, _ = particle_props(np.asarray(pdg)[o])
        p0 = np.sqrt(_col(steps, "px0") ** 2 + _col(steps, "py0") ** 2 + _col(steps, "pz0") ** 2)[o]
        pos = lambda e: np.column_stack([_col(steps, f"{a}{e}").astype(float) for a in "xyz"])[o]  # noqa: E731
        self.cols = {
            "event": torch.from_numpy((ev & 0xFFFFFFFF).astype(np.uint32).view(np.int32)).to("cuda"),
            "bar": _dev(_col(steps, "bar")[o].astype(np.int32), np.int32),
            "t0": _dev(t0[o], np.float64), "t1": _dev(t1[o], np.float64),
            "pos0": _dev(pos(0).ravel(), np.float64), "pos1": _dev(pos(1).ravel(), np.float64),
            "edep": _dev(_col(steps, "edep")[o].astype(np.float64), np.float64),
            "len": _dev(_col(steps, "len")[o].astype(np.float64), np.float64),
            "p0": _dev(p0, np.float64), "p1": _dev(_col(steps, "p1")[o].astype(np.float64), np.float64),
            "mass": _dev(mass, np.float64), "charge": _dev(charge, np.float64)}

    def struct(self):
        c = self.cols
        return L.LookupSteps(*[L.ptr(c[k]) for k in ("event", "bar", "t0", "t1", "pos0", "pos1", "edep", "len",
                                                     "p0", "p1", "mass", "charge")], self.n)


def device_photons(ctx, ls, cap_hint=0):
    """Run crv_lookup_photons_run on LookupSteps: device buffers label, time, photon (row in ls order), j, rank and
    the count, the output grown once to the size the library reports."""
    import ctypes
    import torch
    from .chain import _empty
    cap = max(int(cap_hint), 4096)
    for _ in range(2):
        pb = {k: _empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                             ("photon", torch.int64), ("j", torch.int32), ("rank", torch.int32))}
        po = L.Photons(*[L.ptr(pb[k]) for k in ("label", "time", "photon", "j", "rank")], cap, 0, 0)
        rc = ctx.lib.crv_lookup_photons_run(ctx.h, ctypes.byref(ls.struct()), ctypes.byref(po), None)
        if rc == L.CRV_E_CAPACITY:
            cap = po.n
            continue
        ctx.check(rc)
        return pb, po
    raise RuntimeError("lookup photons: capacity not met after growing")


def make_photons(steps, pdg, geom, cfg, tables, deviation=None, rng=None, key=None, cond=None, evkey=None):
    """Photons at the SiPMs for a CrvStep table (reference response/photons.py:359-371).

    steps: table with bar, edep, t0, t1, x0..z1, px0, py0, pz0, p1, len (Mu2e frame, mm, ns, MeV); pdg: code per
    step; geom: a CRVGeometry; cfg: a ChainConfig (its response.sectors / speed_ratio / atten_ratio and yield scales);
    tables: {name: crvhip.lookup.LookupTable}. The photon-yield deviation per channel comes from `cond` (default: the
    synthetic detector's nominal conditions, deviation 0) unless `deviation` is given. Returns (step index,
    channel = bar * 4 + SiPM, arrival time ns) in (step, SiPM, rank) order; steps with a NaN start or end time give
    none. As in the reference, no channel status applies here (charges() masks status bits 0 and 2)."""
    from . import chain as CH, detector as DT
    from dataclasses import replace
    if cond is None:
        cond = DT.nominal_conditions()
    cond = replace(cond, status=np.zeros(len(np.asarray(cond.status)), np.int64), readout_mask=None)
    if deviation is not None:
        cond = replace(cond, deviation=np.asarray(deviation, np.float64))
    ev = np.asarray(evkey, np.int64) if evkey is not None else np.zeros(len(_col(steps, "bar")), np.int64)
    ctx = CH.context(cfg, cond)
    upload(ctx, tables, geom, cfg.response)
    ls = LookupSteps(ev, steps, pdg)
    ctx.set_params(cfg, CH._seed(key, rng), ls.evkey_hi)
    if ls.n == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    pb, po = device_photons(ctx, ls)
    n = po.n
    lab = CH._host(pb["label"], n)
    return ls.order[CH._host(pb["photon"], n)], lab % CH.CHAN_MULT, CH._host(pb["time"], n)
