"""Photon arrivals at the SiPMs from CrvSteps through the v6 lookup tables (oracle; test infrastructure only).

Restates the reference's response/photons.py:
  particle kinematics      beta_of (:80-85): beta from the mean of the start and end energies, massless -> 1;
  Cerenkov yield           cerenkov_per_mm (:88-92): linear from (0, 0) through the table's map, flat beyond it;
  bin lookup               find_bin (:95-99): the first i with edges[i] <= v <= edges[i + 1], -1 outside;
  photon numbers           _draw_counts / _per_substep (:102-112): Gauss(avg, sqrt avg) if avg > 10 else
                           Poisson(avg), divided by the substep count, rounded half to even, at least 0;
  geometry                 _inside_scint (:139-146), _inside_fiber (:149-160);
  substep loop             _substep_photons (:163-224): position and time at the substep centre, scintillation and
                           scintillator-Cerenkov bins inside the scintillator, else the fibre-Cerenkov bin in a
                           fibre; p = prob * dev [* e^(-mu z)], capped at 1; Binomial(n, p) arrivals; each at the
                           substep time + a delay index drawn from the bin's weights (first j with u * sum <= cum_j);
  rows                     iter_rows (:305-356): SiPM mirroring (1: z -> -z, 2: y -> -y, 3: both; reflector sectors
                           simulate SiPMs 0 and 2, reflector < 0 flips z first and stores SiPM s at s + 1, its
                           deviation still that of s), speed ratio as an extra delay k (z + H) on both endpoint times,
                           attenuation ratio as mu = (1/r - 1) / L_table;
  table propagation        table_propagation (:230-254);
and the stock SiPM-stage mask (chain.py:199-204): channels with status bit 0 or 2 give no photons.

Every draw is keyed by Philox (oracle.philox) on the photon's identity instead of the reference's sequential
Generator: photon numbers per (step, SiPM, type) (KIND_LCOUNT; counter (event, channel + type * 2^15, step ordinal,
call)), Binomial arrivals per substep (KIND_LBINOM; counter (event, channel + call * 2^15, step ordinal, substep):
call 0 holds chunk 0 of the three types, chunk c >= 1 of type t is call 1 + 3 ((c - 1) >> 2) + t, word (c - 1) & 3),
and each arriving photon's delay (KIND_LDELAY; counter (event, channel, step ordinal, rank >> 2), word rank & 3).
A row's photons are ranked substep by substep, then by type, then by arrival. The arithmetic is a fixed sequence
of IEEE double operations (oracle.fmath for exp, log, cos and acos; hypot and vector norms written as square roots of
sums of squares) that the HIP kernel repeats, so the arrivals are bit-identical.
"""
import numpy as np

from . import fmath as F
from . import philox as PX
from . import rand as R

PRECISION_MM = 0.1
CHAN_MULT = 32768
GAUSS_ABOVE = 10.0
MAX_SUBSTEPS = 1 << 24


def find_bin(edges, v):
    edges = np.asarray(edges, np.float64)
    v = np.asarray(v, np.float64)
    i = np.searchsorted(edges, v, side="left")
    i = np.where(v == edges[0], 1, i)
    return np.where((i >= 1) & (i <= len(edges) - 1), i - 1, -1)


def interp_map(beta, cmap):
    """Photons per mm at beta: 0 at beta <= 0, linear between (0, 0) and the map's points, the last value beyond."""
    b = np.r_[0.0, np.asarray(cmap, np.float64)[:, 0]]
    n = np.r_[0.0, np.asarray(cmap, np.float64)[:, 1]]
    beta = np.asarray(beta, np.float64)
    j = np.clip(np.searchsorted(b, beta, side="right") - 1, 0, len(b) - 2)
    mid = n[j] + (beta - b[j]) * ((n[j + 1] - n[j]) / (b[j + 1] - b[j]))
    out = np.where(beta <= b[0], n[0], mid)
    return np.where(beta >= b[-1], n[-1], out)


def beta_of(p0, p1, mass):
    e = 0.5 * (np.sqrt(p0 * p0 + mass * mass) + np.sqrt(p1 * p1 + mass * mass))
    g = e / np.where(mass > 0, mass, 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        b = np.sqrt(np.clip(1.0 - 1.0 / (g * g), 0.0, 1.0))
    return np.where(mass > 0, b, 1.0)


def hyp(a, b):
    return np.sqrt(a * a + b * b)


def inside_scint(L, x, y, z):
    ax, ay = np.abs(x), np.abs(y)
    ok = (ax < L.half_thickness) & (ay < L.half_width) & (np.abs(z) < L.half_length)
    ex = ax / L.hole_radius_x
    ey = (ay - L.fiber_separation / 2) / L.hole_radius_y
    ok &= (ex * ex + ey * ey) >= 1.0
    cx, cy = L.half_thickness - L.corner_radius, L.half_width - L.corner_radius
    corner = (ax > cx) & (ay > cy) & (hyp(ax - cx, ay - cy) > L.corner_radius)
    return ok & ~corner


def inside_fiber(L, x, y, z, dx, dy):
    ok = np.abs(z) < L.half_length
    h = L.fiber_separation / 2
    r0, r1 = hyp(x, y + h), hyp(x, y - h)
    in0, in1 = r0 <= L.fiber_radius, r1 <= L.fiber_radius
    ry = np.where(in1, y - h, y + h)
    r = np.where(in1, r1, r0)
    norm = hyp(dx, dy) * r
    cosang = np.where(norm > 0, (dx * x + dy * ry) / np.where(norm > 0, norm, 1.0), 1.0)
    phi = np.where(norm > 0, F.acos(np.clip(cosang, -1.0, 1.0)), 0.0)
    return ok & (in0 | in1), r, phi


def table_propagation(L):
    """(inverse signal speed [ns/mm], attenuation length [mm]) of a table: slope of the 10 % quantile of the
    scintillation delays against the distance to the SiPM end, and -1 / slope of the log of the summed arrival
    probability, for distances between 0.2 and 1.8 half lengths."""
    E, T = L.edges, L.tables["scint"]
    nx, ny, nz = len(E["x"]) - 1, len(E["y"]) - 1, len(E["z"]) - 1
    p = T.prob.astype(np.float64).reshape(nx, ny, nz)
    ds = T.delay_sum.astype(np.float64).reshape(nx, ny, nz)
    w = np.where(ds > 0, p / np.where(ds > 0, ds, 1.0), 0.0)
    h = np.einsum("xyz,xyzj->zj", w, T.delays.astype(np.float64).reshape(nx, ny, nz, -1))
    d = 0.5 * (E["z"][1:] + E["z"][:-1]) + L.half_length
    tot = h.sum(1)
    ok = tot > 0
    q = np.argmax(np.cumsum(h, 1) >= 0.1 * tot[:, None], axis=1)
    inv_v = np.polyfit(d[ok], q[ok].astype(float), 1)[0]
    P = p.sum((0, 1))
    a = (P > 0) & (d > 0.2 * L.half_length) & (d < 1.8 * L.half_length)
    sl = np.polyfit(d[a], np.log(P[a]), 1)[0]
    if not sl < 0:
        raise ValueError(f"cannot determine the attenuation length of {L.name}")
    return float(inv_v), float(-1.0 / sl)


def _counts(avg, nsub, seed, ev, c1, j):
    """Photons per substep of one type per row: (n int64, capped Poisson draws)."""
    if not np.all(np.isfinite(avg)):
        raise ValueError("non-finite photon-number average")
    big = avg > GAUSS_ABOVE
    x = np.zeros(len(avg))
    capped = 0
    if big.any():
        g = np.flatnonzero(big)
        u = PX.draw(seed, PX.KIND_LCOUNT, ev[g], c1[g], j[g], 0)
        x[g] = avg[g] + np.sqrt(avg[g]) * R.normal(u[0], u[1])
    s = np.flatnonzero(~big)
    if len(s):
        k, cap = R.poisson(np.maximum(avg[s], 0.0), lambda call: PX.draw(seed, PX.KIND_LCOUNT, ev[s], c1[s], j[s], call))
        x[s] = k.astype(np.float64)
        capped = int(cap.sum())
    return np.maximum(np.rint(x / nsub.astype(np.float64)), 0.0).astype(np.int64), capped


def rows(st, bars, tables, props, deviation, status, yield_scale, variation_scale, seed):
    """Per (step, SiPM slot) rows, step-major, slot 0-3, of the active ones: dict of arrays.

    st: evkey (sorted), bar, t0, t1, pos0 / pos1 [n, 3] (world mm), edep, len, p0, p1, mass, charge;
    bars: per bar table (index into tables, -1: not simulated), reflector, yield (photons / MeV), speed_ratio,
    atten_ratio, center [nbars, 3], l2w [nbars, 3]; props: (inv_v, att) per table."""
    ev = np.asarray(st["evkey"], np.int64)
    n = len(ev)
    from .response import step_ordinals
    jst = step_ordinals(ev)
    bar = np.asarray(st["bar"], np.int64)
    tid = np.asarray(bars["table"], np.int64)[bar]
    refl = np.asarray(bars["reflector"], np.int64)[bar]
    l2w = np.asarray(bars["l2w"], np.int64)[bar]
    cen = np.asarray(bars["center"], np.float64)[bar]
    r_ = np.arange(n)
    sloc = (np.asarray(st["pos0"], np.float64) - cen)[r_[:, None], l2w]
    eloc = (np.asarray(st["pos1"], np.float64) - cen)[r_[:, None], l2w]
    t0, t1 = np.asarray(st["t0"], np.float64), np.asarray(st["t1"], np.float64)
    good = np.isfinite(t0) & np.isfinite(t1) & (tid >= 0)
    step = np.repeat(r_, 4)
    s = np.tile(np.arange(4), n)
    rf = refl[step]
    sim = (rf == 0) | (s == 0) | (s == 2)
    chan = bar[step] * 4 + np.where(rf < 0, s + 1, s)
    chdev = bar[step] * 4 + s
    act = good[step] & sim
    act &= (np.asarray(status)[np.where(act, chan, 0)] & 0b101) == 0
    a_ = np.flatnonzero(act)
    step, s, chan, chdev, rf = step[a_], s[a_], chan[a_], chdev[a_], rf[a_]
    a = sloc[step].copy()
    b = eloc[step].copy()
    neg = rf < 0
    a[neg, 2], b[neg, 2] = -a[neg, 2], -b[neg, 2]
    fz = (s == 1) | (s == 3)
    a[fz, 2], b[fz, 2] = -a[fz, 2], -b[fz, 2]
    fy = (s == 2) | (s == 3)
    a[fy, 1], b[fy, 1] = -a[fy, 1], -b[fy, 1]
    d = b - a
    dl = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    theta = np.where(dl > 0, F.acos(np.clip(d[:, 2] / np.where(dl > 0, dl, 1.0), -1.0, 1.0)), 0.0)
    q = dl / PRECISION_MM
    nsub = np.where((q >= 1.0) & (q < 9.2e18), np.floor(np.where((q >= 1.0) & (q < 9.2e18), q, 1.0)), 1.0)
    nsub = nsub.astype(np.int64)
    if np.any(nsub > MAX_SUBSTEPS):
        raise ValueError("step longer than the substep limit")
    t = tid[step]
    H = np.array([tables[i].half_length for i in range(len(tables))])[t]
    inv_v = np.array([props[i][0] for i in range(len(tables))])[t]
    att = np.array([props[i][1] for i in range(len(tables))])[t]
    bstep = bar[step]
    k = inv_v * (1.0 / np.asarray(bars["speed_ratio"], np.float64)[bstep] - 1.0)
    mu = (1.0 / np.asarray(bars["atten_ratio"], np.float64)[bstep] - 1.0) / att
    ta = t0[step] + k * (a[:, 2] + H)
    tb = t1[step] + k * (b[:, 2] + H)
    dev = (np.asarray(deviation, np.float64)[chdev] * variation_scale + 1.0) * yield_scale
    beta = beta_of(np.asarray(st["p0"], np.float64), np.asarray(st["p1"], np.float64),
                   np.asarray(st["mass"], np.float64))[step]
    qq = np.abs(np.asarray(st["charge"], np.float64))[step]
    ln = np.asarray(st["len"], np.float64)[step]
    Y = np.asarray(bars["yield"], np.float64)[bstep]
    evr, jr = ev[step], jst[step]
    ck_s = np.zeros(len(step))
    ck_f = np.zeros(len(step))
    for i, L in enumerate(tables):
        m = t == i
        ck_s[m] = interp_map(beta[m], L.ck_scint) * qq[m]
        ck_f[m] = interp_map(beta[m], L.ck_fiber) * qq[m]
    avgs = (Y * np.asarray(st["edep"], np.float64)[step], ck_s * ln, ck_f * ln)
    ns, capped = [], 0
    for typ, avg in enumerate(avgs):
        nn, c = _counts(avg, nsub, seed, evr, chan + (typ << 15), jr)
        ns.append(nn)
        capped += c
    return {"step": step, "s": s, "chan": chan, "table": t, "a": a, "d": d, "ta": ta, "tb": tb, "nsub": nsub,
            "beta": beta, "theta": theta, "dev": dev, "mu": mu, "n": np.stack(ns, 1), "evkey": evr, "j": jr,
            "capped": capped}


def _bins(L, x, y, z, beta, theta, r, phi, insc, infb):
    E = L.edges
    nb = {k: len(v) - 1 for k, v in E.items()}
    bx, by, bz = find_bin(E["x"], np.abs(x)), find_bin(E["y"], y), find_bin(E["z"], z)
    bb = find_bin(E["beta"], beta)
    geo = (bx >= 0) & (by >= 0) & (bz >= 0)
    sbin = np.where(insc & geo, bz + by * nb["z"] + bx * nb["y"] * nb["z"], -1)
    cbin = np.where(insc & geo & (bb >= 0), bb + bz * nb["beta"] + by * nb["z"] * nb["beta"]
                    + bx * nb["y"] * nb["z"] * nb["beta"], -1)
    bth, bph, br = find_bin(E["theta"], theta), find_bin(E["phi"], phi), find_bin(E["r"], r)
    fok = infb & (bb >= 0) & (bth >= 0) & (bph >= 0) & (br >= 0) & (bz >= 0)
    fbin = np.where(fok, bz + br * nb["z"] + bph * nb["r"] * nb["z"] + bth * nb["phi"] * nb["r"] * nb["z"]
                    + bb * nb["theta"] * nb["phi"] * nb["r"] * nb["z"], -1)
    return sbin, cbin, fbin


def _binom_uniforms(seed, ev, chan, j, sub, typ):
    """uniform_of(chunk) for draws of one type on substeps (ev, chan, j, sub arrays)."""
    def uniform_of(c):
        if c == 0:
            return PX.draw(seed, PX.KIND_LBINOM, ev, chan, j, sub)[typ]
        call = 1 + 3 * ((c - 1) >> 2) + typ
        return PX.draw(seed, PX.KIND_LBINOM, ev, chan + (call << 15), j, sub)[(c - 1) & 3]
    return uniform_of


def photons(st, bars, tables, props, deviation, status, yield_scale, variation_scale, seed):
    """Photon arrivals in (row, rank) order: dict evkey, chan, row (step index), j, rank, time, typ, bin, sub (substep)
    and delay (index drawn from the bin); plus 'capped' (Poisson draws that hit the rejection bound) and the rows."""
    for L in tables:
        if L.fiber_decay_time > 0:
            raise ValueError("fiber_decay_time > 0 (fibre-emission decays) is not supported")
    Rw = rows(st, bars, tables, props, deviation, status, yield_scale, variation_scale, seed)
    nsub = Rw["nsub"]
    nr = len(nsub)
    ri = np.repeat(np.arange(nr), nsub)
    kk = np.arange(len(ri)) - np.repeat(np.cumsum(nsub) - nsub, nsub)
    frac = (kk.astype(np.float64) + 0.5) / nsub[ri].astype(np.float64)
    p = Rw["a"][ri] + Rw["d"][ri] * frac[:, None]
    tsub = Rw["ta"][ri] + (Rw["tb"][ri] - Rw["ta"][ri]) * frac
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    narr = np.zeros((len(ri), 3), np.int64)
    bins = np.full((len(ri), 3), -1, np.int64)
    for i, L in enumerate(tables):
        m = np.flatnonzero(Rw["table"][ri] == i)
        if not len(m):
            continue
        insc = inside_scint(L, x[m], y[m], z[m])
        infb, r, phi = inside_fiber(L, x[m], y[m], z[m], Rw["d"][ri[m], 0], Rw["d"][ri[m], 1])
        infb &= ~insc
        bb = _bins(L, x[m], y[m], z[m], Rw["beta"][ri[m]], Rw["theta"][ri[m]], r, phi, insc, infb)
        for typ, tname in enumerate(("scint", "ckscint", "ckfiber")):
            T = L.tables[tname]
            bins[m, typ] = bb[typ]
            nn = Rw["n"][ri[m], typ]
            g = np.flatnonzero((bb[typ] >= 0) & (nn > 0))
            if not len(g):
                continue
            rows_g = ri[m[g]]
            pr = T.prob[bb[typ][g]].astype(np.float64) * Rw["dev"][rows_g]
            mu = Rw["mu"][rows_g]
            pr = np.where(mu != 0.0, pr * F.exp(-mu * z[m[g]]), pr)
            pr = np.minimum(pr, 1.0)
            uo = _binom_uniforms(seed, Rw["evkey"][rows_g], Rw["chan"][rows_g], Rw["j"][rows_g], kk[m[g]], typ)
            kb, ok = R.binomial(nn[g], pr, uo)
            if not ok.all():
                raise ValueError("Binomial draw beyond the draw layout")
            narr[m[g], typ] = kb
    # ranks: substep by substep, then type, then arrival
    per = narr.ravel()
    tot = int(per.sum())
    rowcount = np.bincount(ri, weights=narr.sum(1), minlength=nr).astype(np.int64)
    if np.any(rowcount > 0xFFFF):
        raise ValueError("more than 65535 photons from one (step, SiPM)")
    sub_of = np.repeat(np.arange(len(per)), per)
    si, typ = sub_of // 3, sub_of % 3
    row = ri[si]
    rank = np.arange(tot) - np.repeat((np.cumsum(rowcount) - rowcount), rowcount)
    ev, chan, j = Rw["evkey"][row], Rw["chan"][row], Rw["j"][row]
    u = np.zeros(tot)
    for w in range(4):
        m = np.flatnonzero((rank & 3) == w)
        if len(m):
            u[m] = PX.draw(seed, PX.KIND_LDELAY, ev[m], chan[m], j[m], rank[m] >> 2)[w]
    bsel = bins[si, typ]
    time = np.empty(tot)
    delay = np.zeros(tot, np.int64)
    for i, L in enumerate(tables):
        for t_, tname in enumerate(("scint", "ckscint", "ckfiber")):
            m = np.flatnonzero((Rw["table"][row] == i) & (typ == t_))
            if not len(m):
                continue
            T = L.tables[tname]
            cum = np.cumsum(T.delays[bsel[m]].astype(np.int64), axis=1)
            thr = u[m] * T.delay_sum[bsel[m]].astype(np.float64)
            delay[m] = (cum.astype(np.float64) < thr[:, None]).sum(1)
            time[m] = tsub[si[m]] + delay[m].astype(np.float64)
    return {"evkey": ev, "chan": chan, "row": Rw["step"][row], "j": j, "rank": rank, "time": time, "typ": typ,
            "bin": bsel, "sub": kk[si], "delay": delay, "capped": Rw["capped"], "rows": Rw}
