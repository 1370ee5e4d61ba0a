"""CRV coincidence-finder oracle (test infrastructure only).

Restates the reference's reco/coinc.py:26-392 (SectorCuts, FinderConfig, _type_limits, _find_clusters,
_check_combination, _check_coincidence, find_event, _properties) and reco/coinc_cpu.py:11-51 (the opposite-side
attachment of the default CPU mode). Per-cluster sums (PE, PE-weighted position, slope moments, side sums) run
sequentially in member order, which the HIP kernel repeats; the reference's numpy sums may group terms
pairwise, so the cluster moments agree with it to rounding, the memberships exactly.
"""
from dataclasses import dataclass, field
from itertools import combinations, chain, product

import numpy as np

N_LAYERS = 4


@dataclass(frozen=True)
class SectorCuts:
    pe_threshold: float
    max_dt_adjacent: float
    max_dt: float
    min_slope: float
    max_slope: float
    max_slope_diff: float
    coincidence_layers: int
    min_cluster_pes: float
    initial_max_distance: float
    final_max_distance: float


@dataclass(frozen=True)
class FinderConfig:
    fiber_signal_speed: float = 140.0
    max_time_interval: float = 20.0
    time_offset: float = 34.5
    compensate_status_bits: tuple = (0, 1, 2)
    ignore_secondary_pulses: bool = True
    big_cluster_threshold: int = 50000
    cuts: dict = field(default_factory=dict)
    add_opposite_side_pulses: bool = True


def _ssum(vals):
    s = 0.0
    for v in vals:
        s = s + v
    return s


def type_limits(G, cfg):
    out = {}
    for t in np.unique(G.sector_type):
        secs = np.flatnonzero(G.sector_type == t)
        cuts = [cfg.cuts[G.sector_names[s]] for s in secs if G.sector_names[s] in cfg.cuts]
        if not cuts:
            continue
        p0, p1 = G.pos_side0[secs].min(), G.pos_side1[secs].max()
        out[int(t)] = dict(p0=p0, p1=p1, half=(p1 - p0) / 2, max_dt=max(c.max_dt for c in cuts),
                           side0=bool(G.sipms_side0[secs].any()), side1=bool(G.sipms_side1[secs].any()))
    return out


def find_clusters(hits, max_dt, opposite=False, dt_opp=0.0):
    hits = list(hits)
    clusters = []
    while hits:
        cluster = []
        last = -1
        while last != len(cluster):
            last = len(cluster)
            i = 0
            while i < len(hits):
                h = hits[i]
                if not cluster:
                    cluster.append(h)
                    hits.pop(i)
                    continue
                joined = False
                for c in cluster:
                    md = max(h["maxd"], c["maxd"])
                    mt = max_dt + (dt_opp if opposite and h["side"] != c["side"] else 0.0)
                    if abs(h["x"] - c["x"]) <= md and abs(h["t"] - c["t"]) < mt:
                        cluster.append(h)
                        hits.pop(i)
                        joined = True
                        break
                if not joined:
                    i += 1
        clusters.append(cluster)
    return clusters


def check_combination(combo):
    max_dt = max(h["cuts"].max_dt for h in combo)
    ts = [h["t"] for h in combo]
    if max(ts) - min(ts) > max_dt:
        return False
    lo = min(h["cuts"].min_slope for h in combo)
    hi = max(h["cuts"].max_slope for h in combo)
    msd = max(h["cuts"].max_slope_diff for h in combo)
    slopes = []
    for a, b in zip(combo[:-1], combo[1:]):
        s = (b["x"] - a["x"]) / (b["y"] - a["y"])
        if s < lo or s > hi:
            return False
        slopes.append(s)
    n = len(combo)
    if n > 2 and abs(slopes[0] - slopes[1]) > msd:
        return False
    if n > 3 and (abs(slopes[0] - slopes[2]) > msd or abs(slopes[1] - slopes[2]) > msd):
        return False
    srt = sorted(combo, key=lambda h: h["x"])
    for i, h in enumerate(srt):
        if i > 0:
            h["maxd"] = max(h["maxd"], h["x"] - srt[i - 1]["x"])
        if i + 1 < n:
            h["maxd"] = max(h["maxd"], srt[i + 1]["x"] - h["x"])
    return True


def check_coincidence(cluster, cfg):
    if not cluster:
        return []
    by_layer = [[] for _ in range(N_LAYERS)]
    for h in cluster:
        hh = dict(h)
        hh["maxd"] = h["cuts"].final_max_distance
        by_layer[h["layer"]].append(hh)
    nmin = min(h["cuts"].coincidence_layers for h in cluster)
    nmax = max(h["cuts"].coincidence_layers for h in cluster)
    if len(cluster) > cfg.big_cluster_threshold and sum(1 for l in by_layer if l) >= nmin:
        return list(cluster)
    sizes = [len(l) for l in by_layer]
    total = sum(int(np.prod([sizes[l] for l in layers]))
                for n, need in ((2, nmin == 2), (3, nmin <= 3 <= nmax), (4, nmax == 4)) if need
                for layers in combinations(range(N_LAYERS), n))
    if total > NUMPY_COMBOS:
        return _check_coincidence_np(by_layer, nmin, nmax)
    keep = {}
    for n, need in ((2, nmin == 2), (3, nmin <= 3 <= nmax), (4, nmax == 4)):
        if not need:
            continue
        for layers in combinations(range(N_LAYERS), n):
            for combo in product(*(by_layer[l] for l in layers)):
                if all(h["cuts"].coincidence_layers > n for h in combo):
                    continue
                if check_combination(list(combo)):
                    for h in combo:
                        keep[h["idx"]] = h
    return [keep[k] for k in sorted(keep)]


NUMPY_COMBOS = 2000  # clusters with more layer combinations than this take the numpy path


def _check_coincidence_np(by_layer, nmin, nmax, chunk=1 << 20):
    """check_coincidence over numpy arrays of the combinations, chunked, with every hit's own cuts: the same tests
    as check_combination (time window with the combination's widest max_dt, slopes within its widest slope range,
    slope differences, the skip of combinations whose hits all need more layers) and the same widening (x gaps to the
    neighbours of a stable sort by x of the layer-ordered combination), as the reference's numpy variant
    reco/coinc.py:209-250 does for uniform cuts. Times are finite."""
    H = [h for l in by_layer for h in l]
    off = np.cumsum([0] + [len(l) for l in by_layer])
    T = np.array([h["t"] for h in H], float)
    X = np.array([h["x"] for h in H], float)
    Y = np.array([h["y"] for h in H], float)
    c = [h["cuts"] for h in H]
    CL = np.array([q.coincidence_layers for q in c])
    MDT = np.array([q.max_dt for q in c], float)
    LO = np.array([q.min_slope for q in c], float)
    HI = np.array([q.max_slope for q in c], float)
    MSD = np.array([q.max_slope_diff for q in c], float)
    wid = np.full(len(H), -np.inf)
    kept = np.zeros(len(H), bool)
    for n, need in ((2, nmin == 2), (3, nmin <= 3 <= nmax), (4, nmax == 4)):
        if not need:
            continue
        for layers in combinations(range(N_LAYERS), n):
            shape = tuple(int(off[l + 1] - off[l]) for l in layers)
            P = int(np.prod(shape))
            for a in range(0, P, chunk):
                idx = np.unravel_index(np.arange(a, min(a + chunk, P)), shape)
                cb = np.stack([off[l] + i for l, i in zip(layers, idx)], 1)
                t, x, y = T[cb], X[cb], Y[cb]
                ok = ~(CL[cb] > n).all(1)
                ok &= ~(t.max(1) - t.min(1) > MDT[cb].max(1))
                with np.errstate(divide="ignore", invalid="ignore"):
                    sl = (x[:, 1:] - x[:, :-1]) / (y[:, 1:] - y[:, :-1])
                lo, hi, msd = LO[cb].min(1), HI[cb].max(1), MSD[cb].max(1)
                ok &= ~((sl < lo[:, None]) | (sl > hi[:, None])).any(1)
                if n > 2:
                    ok &= ~(np.abs(sl[:, 0] - sl[:, 1]) > msd)
                if n > 3:
                    ok &= ~((np.abs(sl[:, 0] - sl[:, 2]) > msd) | (np.abs(sl[:, 1] - sl[:, 2]) > msd))
                cb, x = cb[ok], x[ok]
                if not len(cb):
                    continue
                o = np.argsort(x, axis=1, kind="stable")
                xs, hs = np.take_along_axis(x, o, 1), np.take_along_axis(cb, o, 1)
                gap = np.diff(xs, axis=1)
                w = np.full(xs.shape, -np.inf)
                w[:, 1:] = gap
                w[:, :-1] = np.maximum(w[:, :-1], gap)
                np.maximum.at(wid, hs.ravel(), w.ravel())
                kept[hs.ravel()] = True
    out = []
    for i in np.flatnonzero(kept):
        h = H[i]
        h["maxd"] = max(h["maxd"], float(wid[i]))
        out.append(h)
    return sorted(out, key=lambda h: h["idx"])


def attach_opposite_side_pulses(clusters, hits, max_dt):
    """Reference reco/coinc_cpu.py:11-51 (member snapshot, nearest cluster per candidate, one pulse per SiPM)."""
    if not clusters or not hits:
        return
    members = list(chain.from_iterable(clusters))
    if not members:
        return
    fields = ("idx", "bar", "sipm", "side", "t", "seq")
    m = {k: np.fromiter((h[k] for h in members), float if k == "t" else np.int64) for k in fields}
    h = {k: np.fromiter((hit[k] for hit in hits), float if k == "t" else np.int64) for k in fields}
    mc = np.repeat(np.arange(len(clusters)), [len(cl) for cl in clusters])
    candidates = np.flatnonzero((h["seq"] == 0) & ~np.isin(h["idx"], m["idx"]))
    anchors = np.flatnonzero(m["seq"] == 0)
    if not len(candidates) or not len(anchors):
        return
    occupied = set(zip(mc.tolist(), m["bar"].tolist(), m["side"].tolist()))
    c = candidates
    dt = np.abs(h["t"][c, None] - m["t"][anchors])
    matches = ((h["bar"][c, None] == m["bar"][anchors]) &
               (h["side"][c, None] == 1 - m["side"][anchors]) & (dt < max_dt))
    row, col = np.nonzero(matches)
    candidate, cluster = c[row], mc[anchors[col]]
    delta = dt[row, col]
    available = np.array([(int(a), int(b), int(s)) not in occupied
                          for a, b, s in zip(cluster, h["bar"][candidate], h["side"][candidate])], bool)
    candidate, cluster, delta = candidate[available], cluster[available], delta[available]
    order = np.lexsort((cluster, delta, candidate))
    candidate, cluster, delta = candidate[order], cluster[order], delta[order]
    first = np.r_[True, candidate[1:] != candidate[:-1]] if len(candidate) else np.zeros(0, bool)
    candidate, cluster, delta = candidate[first], cluster[first], delta[first]
    order = np.lexsort((candidate, delta, h["sipm"][candidate], h["bar"][candidate], cluster))
    candidate, cluster = candidate[order], cluster[order]
    keys = list(zip(cluster.tolist(), h["bar"][candidate].tolist(), h["sipm"][candidate].tolist()))
    seen = set()
    for k, cl, cd in zip(keys, cluster, candidate):
        if k in seen:
            continue
        seen.add(k)
        clusters[cl].append(hits[cd])


def find_event(pulses, G, status, cfg):
    """pulses: dict of arrays for ONE event in collection order (idx, bar, sipm, time, pe, seq)."""
    lim = type_limits(G, cfg)
    hits = {}
    for i in range(len(pulses["idx"])):
        if cfg.ignore_secondary_pulses and pulses["seq"][i] > 0:
            continue
        bar = int(pulses["bar"][i])
        s = int(G.sector[bar])
        name = G.sector_names[s]
        if name not in cfg.cuts:
            continue
        t_type = int(G.sector_type[s])
        sipm = int(pulses["sipm"][i])
        side = sipm % 2
        t = float(pulses["time"][i])
        if side == 0:
            t += (G.pos_side0[s] - lim[t_type]["p0"]) / cfg.fiber_signal_speed
        else:
            t += (lim[t_type]["p1"] - G.pos_side1[s]) / cfg.fiber_signal_speed
        pe = float(pulses["pe"][i])
        other = bar * 4 + sipm + (2 if sipm < 2 else -2)
        if any((int(status[other]) >> b) & 1 for b in cfg.compensate_status_bits):
            pe *= 2.0
        cuts = cfg.cuts[name]
        hits.setdefault(t_type, []).append(dict(idx=int(pulses["idx"][i]), bar=bar, sector=s, layer=int(G.layer[bar]),
                                               counter=int(G.counter_in_type[bar]), side=side, sipm=sipm,
                                               seq=int(pulses["seq"][i]), x=float(G.x[bar]),
                                               y=float(G.y[bar]), t=t, pe=pe, cuts=cuts,
                                               maxd=cuts.initial_max_distance))
    out = []
    for t_type in sorted(hits):
        H = hits[t_type]
        L = lim[t_type]
        coinc = []
        filtered = []
        for side in (0, 1):
            filt = []
            for h in H:
                if h["side"] != side:
                    continue
                same = adj1 = adj2 = 0.0
                for a in H:
                    if a["side"] != side or a["layer"] != h["layer"] or abs(a["t"] - h["t"]) > h["cuts"].max_dt_adjacent:
                        continue
                    d = a["counter"] - h["counter"]
                    same += a["pe"] if d == 0 else 0.0
                    adj1 += a["pe"] if d == -1 else 0.0
                    adj2 += a["pe"] if d == 1 else 0.0
                if same + adj1 >= h["cuts"].pe_threshold or same + adj2 >= h["cuts"].pe_threshold:
                    filt.append(h)
            filtered.extend(filt)
            for cl in find_clusters(filt, L["max_dt"]):
                coinc += check_coincidence(cl, cfg)
        both = L["side0"] and L["side1"]
        dt_opp = 2.0 * L["half"] / cfg.fiber_signal_speed
        clusters = find_clusters(coinc, L["max_dt"], both, dt_opp)
        if both and cfg.add_opposite_side_pulses:
            attach_opposite_side_pulses(clusters, filtered, L["max_dt"] + dt_opp)
        for cl in clusters:
            row = properties(t_type, cl, G, L, cfg)
            if row is not None:
                out.append(row)
    return out


def properties(t_type, cl, G, L, cfg):
    pe = [h["pe"] for h in cl]
    tot = _ssum(pe)
    if tot < min(h["cuts"].min_cluster_pes for h in cl):
        return None
    x = [h["x"] for h in cl]
    y = [h["y"] for h in cl]
    pos = np.array([_ssum(G.pos[h["bar"]][a] * h["pe"] for h in cl) / tot for a in range(3)])
    sxy = _ssum(pe[i] * x[i] * y[i] for i in range(len(cl)))
    sx = _ssum(pe[i] * x[i] for i in range(len(cl)))
    sy = _ssum(pe[i] * y[i] for i in range(len(cl)))
    syy = _ssum(pe[i] * y[i] * y[i] for i in range(len(cl)))
    den = tot * syy - sy * sy
    slope = (tot * sxy - sx * sy) / den if den != 0 else (np.nan if tot * sxy - sx * sy == 0 else
                                                          np.copysign(np.inf, tot * sxy - sx * sy))
    t = [h["t"] for h in cl]
    side = [h["side"] for h in cl]
    nh, spe, st, tmin, tmax = [], [], [], [], []
    for s in (0, 1):
        sel = [i for i in range(len(cl)) if side[i] == s and pe[i] > 0]
        nh.append(len(sel))
        ps = _ssum(pe[i] for i in sel)
        spe.append(float(ps))
        st.append(float(_ssum(t[i] * pe[i] for i in sel) / ps) if ps > 0 else 0.0)
        tmin.append(min(t[i] for i in sel) if sel else 0.0)
        tmax.append(max(t[i] for i in sel) if sel else 0.0)
    ld = int(G.length_dir[cl[0]["sector"]])
    calc = False
    if nh[0] and nh[1]:
        org = 0.5 * (cfg.fiber_signal_speed * (st[0] - st[1]) + (L["p0"] + L["p1"]))
        if L["p0"] < org < L["p1"] and tmax[0] - tmin[0] < cfg.max_time_interval and tmax[1] - tmin[1] < cfg.max_time_interval:
            calc = True
            pos[ld] = org
            avg_t = st[0] - (org - L["p0"]) / cfg.fiber_signal_speed
        else:
            org = pos[ld]
            avg_t = 0.5 * ((st[0] - (org - L["p0"]) / cfg.fiber_signal_speed) +
                           (st[1] - (L["p1"] - org) / cfg.fiber_signal_speed))
    else:
        org = pos[ld]
        avg_t = (st[0] - (org - L["p0"]) / cfg.fiber_signal_speed) if nh[0] else \
            (st[1] - (L["p1"] - org) / cfg.fiber_signal_speed)
    return dict(sectortype=t_type, start=float(min(t)), end=float(max(t)), pe=float(tot), slope=float(slope),
                layers=sorted({h["layer"] for h in cl}), pulses=[h["idx"] for h in cl], poscalc=calc,
                avgtime=float(avg_t - cfg.time_offset), pos=pos, sidepe=spe, sidetime=st, sidehits=nh)
