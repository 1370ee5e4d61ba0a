"""CRV response oracle: steps -> photoelectrons at the SiPMs -> SiPM charges (test infrastructure only).

Restates, under the closed-form optics of the synthetic detector, the reference's
  response/photons.py:102-112 (_draw_counts / _per_substep: one count draw per (step, SiPM)),
  response/photons.py:163-224 (_substep_photons: arrival time = step time + propagation + emission delay),
  response/sipm.py:80-99 (draw_pixels, _neighbours), response/sipm.py:102-198 (simulate, _simulate_chunk),
  chain.py:199-204 (status mask (status & 0b101) == 0, labels = event key * CHAN_MULT + channel).

Optics (the lookup-table bins of the reference replaced by the closed form of the synthetic detector): for a step
at longitudinal position z on counter c, SiPM s reads out at side s % 2 (z = -H or +H), d its distance to that end,
  mu = edep * yscale * A(d) * dev[channel],  A(d) = w_s e^{-d/l_s} + w_l e^{-d/l_l},  yscale = Y / A(d_ref),
N ~ Poisson(mu) photoelectrons, each arriving at t + d / v_fibre + an Exp(tau) emission delay (the delays of one
(step, SiPM) drawn in ascending order, oracle.rand.ordered_delays).

The SiPM simulation follows the reference loop structure (k-th photon of every SiPM at once, cross-talk cascade
breadth-first, SiPMs ordered by photon count) so that a sequential uniform source in the reference's draw order
reproduces the reference's `sipm.simulate` exactly; the Philox source keys every draw by photon identity
(event key, channel, step ordinal, photon rank, cascade generation, pixel) instead.
"""
import numpy as np

from . import fmath as F
from . import philox as PX
from . import rand as R

NPIX = 41
CHAN_MULT = 32768


# ---------------------------------------------------------------------------------------------------------------------
# photons

def step_ordinals(ev):
    """Index of each step among the steps of its event, in table order."""
    ev = np.asarray(ev)
    o = np.argsort(ev, kind="stable")
    es = ev[o]
    first = np.r_[0, np.flatnonzero(es[1:] != es[:-1]) + 1] if len(es) else np.zeros(0, np.int64)
    start = np.repeat(first, np.diff(np.r_[first, len(es)]))
    j = np.empty(len(ev), np.int64)
    j[o] = np.arange(len(ev)) - start
    return j


def attenuation(d, opt):
    return (opt.w_short * F.exp(-d / opt.l_short)) + (opt.w_long * F.exp(-d / opt.l_long))


def yield_scale(opt):
    return opt.pe_per_mev / attenuation(np.float64(opt.ref_distance), opt)


def photons(steps, status, dev, opt, seed):
    """Photoelectrons at the SiPMs. steps: dict with evkey, counter, edep, z, t (arrays). status, dev: per channel.

    Returns dict of per-photon arrays in (step row, SiPM, rank) order: evkey, chan, row, j, rank, time; plus
    'capped' (number of Poisson draws that hit the rejection bound).
    """
    ev = np.asarray(steps["evkey"], np.int64)
    n = len(ev)
    j = step_ordinals(ev)
    row = np.repeat(np.arange(n), 4)
    s = np.tile(np.arange(4), n)
    ch = 4 * np.asarray(steps["counter"], np.int64)[row] + s
    live = (np.asarray(status)[ch] & 0b101) == 0
    row, s, ch = row[live], s[live], ch[live]
    evr, jr = ev[row], j[row]
    z = np.asarray(steps["z"], np.float64)[row]
    t = np.asarray(steps["t"], np.float64)[row]
    edep = np.asarray(steps["edep"], np.float64)[row]
    d = np.where((s & 1) == 0, z + opt.half_length, opt.half_length - z)
    mu = ((edep * yield_scale(opt)) * attenuation(d, opt)) * np.asarray(dev, np.float64)[ch]
    N, capped = R.poisson(mu, lambda call: PX.draw(seed, PX.KIND_COUNT, evr, ch, jr, call))
    tb = t + d * opt.inv_speed
    # ordered delays, vectorised over groups: rank r advances every group with N > r
    tot = int(N.sum())
    start = np.r_[0, np.cumsum(N)[:-1]]
    time = np.empty(tot)
    S = np.zeros(len(N))
    cache = None
    for r in range(int(N.max(initial=0))):
        g = np.flatnonzero(N > r)
        if r & 3 == 0:
            cache = PX.draw(seed, PX.KIND_DELAY, evr[g], ch[g], jr[g], r >> 2)
            cache_g = g
        u = cache[r & 3][np.searchsorted(cache_g, g)]
        zz = -F.log(u)
        S[g] = S[g] + zz / (N[g] - r).astype(np.float64)
        time[start[g] + r] = tb[g] + opt.tau * S[g]
    rep = lambda a: np.repeat(a, N)
    rank = np.arange(tot) - np.repeat(start, N)
    return {"evkey": rep(evr), "chan": rep(ch), "row": rep(row), "j": rep(jr), "rank": rank, "time": time,
            "capped": int(capped.sum()), "mu": mu, "n": N}


# ---------------------------------------------------------------------------------------------------------------------
# uniform sources for the SiPM simulation

class PhiloxSource:
    """Draws keyed by photon identity: ids = (evkey, chan, j, rank) arrays per photon (input order). Dark discharges
    take the identity (evkey, chan, DARK_J, k), k their index within the SiPM, appended by `thermal`."""

    def __init__(self, seed, ids):
        self.seed = seed
        self.ev, self.ch, self.j, self.r = (np.asarray(a, np.int64) for a in ids)
        self._root = None
        self._dark = None

    def subset(self, keep):
        return PhiloxSource(self.seed, (self.ev[keep], self.ch[keep], self.j[keep], self.r[keep]))

    def _key(self, gid, gen, pix):
        return self.ev[gid], self.ch[gid], self.j[gid], PX.pack_photon(self.r[gid], gen, pix)

    def thermal(self, labels, lam, t_start, t_end):
        """A Poisson(lam) count of dark discharges per SiPM label (the photon-count sampler on the KIND_THERMAL_N
        draws (evkey, chan, 0, call)), discharge k at t_start + (t_end - t_start) u (KIND_THERMAL draw
        (evkey, chan, k, 0), x word)."""
        labels = np.asarray(labels, np.int64)
        ev, ch = labels // CHAN_MULT, labels % CHAN_MULT
        n, capped = R.poisson(np.full(len(labels), lam, np.float64),
                              lambda call: PX.draw(self.seed, PX.KIND_THERMAL_N, ev, ch, 0, call))
        if capped.any() or (n > 0xFFFF).any():
            raise RuntimeError("dark-noise count beyond the draw layout (65535 per SiPM)")
        rep = np.repeat(np.arange(len(labels)), n)
        k = np.arange(len(rep)) - np.repeat(np.cumsum(n) - n, n)
        u = PX.draw(self.seed, PX.KIND_THERMAL, ev[rep], ch[rep], k, 0)
        self._dark = (u[1], u[2])
        self.ev, self.ch = np.r_[self.ev, ev[rep]], np.r_[self.ch, ch[rep]]
        self.j, self.r = np.r_[self.j, np.full(len(rep), PX.DARK_J, np.int64)], np.r_[self.r, k]
        return labels[rep], t_start + (t_end - t_start) * u[0]

    def pixels(self, n):
        u = PX.draw(self.seed, PX.KIND_PIXEL, self.ev, self.ch, self.j, PX.pack_photon(self.r))
        self._root = u[3]
        return u[0], u[1], u[2]

    def dark_pixels(self, n, nx, ny):
        """Uniform pixel of each dark discharge: floor(u nx), floor(u ny) of its KIND_THERMAL draw (y, z words)."""
        ux, uy = self._dark
        return np.floor(ux * nx).astype(np.int64), np.floor(uy * ny).astype(np.int64)

    def fire(self, gid, gen, pix):
        if gen == 0:
            return self._root[gid]
        return PX.draw(self.seed, PX.KIND_XTFIRE, *self._key(gid, gen, pix))[0]

    def xtalk(self, gid, gen, pix):
        return np.stack(PX.draw(self.seed, PX.KIND_XTALK, *self._key(gid, gen, pix)), 1)

    def traps(self, gid, gen, pix, qs):
        """Per trap type: (captured mask, lifetime uniforms of the captured) from the firing's KIND_TRAP draw."""
        u = PX.draw(self.seed, PX.KIND_TRAP, *self._key(gid, gen, pix))
        out = []
        for i, q in enumerate(qs):
            tr = u[2 * i] < q
            out.append((tr, u[2 * i + 1][tr]))
        return out

    def after(self, gid, gen, pix, typ):
        return PX.draw(self.seed, PX.KIND_AFTER, *self._key(np.array([gid]), gen, pix))[typ][0]


class SequentialSource:
    """Uniforms in the reference's draw order from a numpy Generator (pins the restatement to sipm.simulate)."""

    def __init__(self, rng):
        self.rng = rng

    def subset(self, keep):
        return self

    def thermal(self, labels, lam, t_start, t_end):
        n = self.rng.poisson(lam, len(labels))
        tl = np.repeat(np.asarray(labels, np.int64), n)
        return tl, t_start + (t_end - t_start) * self.rng.random(len(tl))

    def pixels(self, n):
        return self.rng.random(n), self.rng.random(n), self.rng.random(n)

    def dark_pixels(self, n, nx, ny):
        return self.rng.integers(0, nx, n), self.rng.integers(0, ny, n)

    def fire(self, gid, gen, pix):
        return self.rng.random(len(gid))

    def xtalk(self, gid, gen, pix):
        return self.rng.random((len(gid), 4))

    def traps(self, gid, gen, pix, qs):
        out = []
        for q in qs:
            tr = self.rng.random(len(gid)) < q
            out.append((tr, self.rng.random(int(tr.sum()))))
        return out

    def after(self, gid, gen, pix, typ):
        return self.rng.random()


# ---------------------------------------------------------------------------------------------------------------------
# SiPM

def p_fire(cfg, v):
    return cfg.p1 * (1.0 - F.exp(-v / cfg.p2))


def q_crosstalk(cfg):
    """Per-neighbour cross-talk candidate probability (1 - (1 - P_ct)^(1/4)) / p_fire(Vov), with ^(1/4) as two
    square roots (reference response/sipm.py:144)."""
    return (1.0 - np.sqrt(np.sqrt(1.0 - cfg.cross_talk))) / p_fire(cfg, np.float64(cfg.overvoltage))


def cdf_of_map(pmap):
    return np.cumsum(np.asarray(pmap, np.float64).ravel())


def pixel_ids(cum, ncol, ub, uhx, uhy, mode):
    """Reference draw_pixels (response/sipm.py:80-91) on given uniforms: bin by CDF, half-pixel bits for lrint."""
    b = np.searchsorted(cum, ub * cum[-1], side="right")
    bx, by = b // ncol, b % ncol
    if mode == "truncate":
        return bx, by
    return bx + (uhx >= 0.5), by + (uhy >= 0.5)


def _pow_int(q, m):
    out = q.copy()
    for i in range(2, int(m.max(initial=1)) + 1):
        out = np.where(m >= i, out * q, out)
    return out


MAX_GEN = 31


def sipm_simulate(label, time, photon, pmap, cfg, source, t_end=1e5, t_start=0.0, thermal_labels=None):
    """SiPM charges for photon arrivals (reference response/sipm.py:102-198).

    label [P] SiPM label per photon, time [P], photon [P] carried id (step row); source: PhiloxSource built on the
    same P photons (in this order) or SequentialSource. thermal_labels: SiPMs that get dark noise at
    cfg.thermal_rate over [t_start, t_end) (sipm.py:113-121: Poisson(rate / p_fire(Vov) * (t_end - t_start))
    discharges per SiPM at uniform times and pixels, photon id -1, merged into the photons' time order).
    cfg.trap0_prob / trap1_prob: each firing is captured by trap i with probability prob_i / p_fire(Vov) and
    releases an afterpulse on its pixel after an Exp(lifetime_i) delay (sipm.py:140, :164-167); the afterpulses of a
    SiPM are tried after all its photons, in time order, with the pixel's recovery at that point and no cross-talk
    (sipm.py:179-193). Returns dict chan, time, charge, pe, photon sorted by (chan, time), record order on ties.
    """
    label = np.asarray(label, np.int64)
    time = np.asarray(time, np.float64)
    photon = np.asarray(photon, np.int64)
    keep = time <= t_end
    label, time, photon = label[keep], time[keep], photon[keep]
    source = source.subset(keep)
    pf_full = p_fire(cfg, np.float64(cfg.overvoltage))
    ndark = 0
    if cfg.thermal_rate > 0 and thermal_labels is not None:
        lam = cfg.thermal_rate / pf_full * (t_end - t_start)
        tl, tt = source.thermal(thermal_labels, lam, t_start, t_end)
        ndark = len(tl)
        label, time = np.r_[label, tl], np.r_[time, tt]
        photon = np.r_[photon, np.full(ndark, -1, np.int64)]
    n = len(label)
    cum = cdf_of_map(pmap)
    ub, uhx, uhy = source.pixels(n)
    px, py = pixel_ids(cum, np.shape(pmap)[1], ub, uhx, uhy, cfg.pixel_id)
    if ndark:
        dx, dy = source.dark_pixels(ndark, cfg.n_pixels_x, cfg.n_pixels_y)
        px, py = np.r_[px[:n - ndark], dx], np.r_[py[:n - ndark], dy]
    dead = np.zeros((NPIX, NPIX), bool)
    for x, y in cfg.inactive:
        dead[x, y] = True
    dead = dead.ravel()
    order = np.lexsort((time, label))
    gid = np.arange(n)[order]
    label, time, photon, px, py = label[order], time[order], photon[order], px[order], py[order]
    empty = {k: np.zeros(0, np.int64 if k in ("chan", "photon") else np.float64)
             for k in ("chan", "time", "charge", "pe", "photon")}
    if n == 0:
        return empty
    u, start, cnt = np.unique(label, return_index=True, return_counts=True)
    by_cnt = np.argsort(-cnt, kind="stable")
    chans, start, cnt = u[by_cnt], start[by_cnt], cnt[by_cnt]
    S = len(chans)
    last = np.full((S, NPIX * NPIX), -np.inf)
    q_ct = q_crosstalk(cfg)
    q_trap = (cfg.trap0_prob / pf_full, cfg.trap1_prob / pf_full)
    life = (cfg.trap0_lifetime, cfg.trap1_lifetime)
    rec = {k: [] for k in ("slot", "time", "v", "photon")}
    traps = []  # per capture: slot, pixel, release time, photon, and the firing's identity (gid, gen) and trap type
    for k in range(int(cnt.max())):
        nn = int((cnt > k).sum())
        idx = start[:nn] + k
        slot = np.arange(nn)
        pix = px[idx] * NPIX + py[idx]
        front = (slot, pix, time[idx], photon[idx], np.ones(nn, np.int64), gid[idx])
        gen = 0
        while len(front[0]):
            sl, pix, t, ph, m, g = front
            alive = ~dead[pix]
            sl, pix, t, ph, m, g = sl[alive], pix[alive], t[alive], ph[alive], m[alive], g[alive]
            v = cfg.overvoltage * (1.0 - F.exp(-(t - last[sl, pix]) / cfg.time_constant))
            p = 1.0 - _pow_int(1.0 - p_fire(cfg, v), m)
            uf = source.fire(g, gen, pix)
            fire = uf < p
            sl, pix, t, ph, v, g = sl[fire], pix[fire], t[fire], ph[fire], v[fire], g[fire]
            last[sl, pix] = t
            rec["slot"].append(sl), rec["time"].append(t), rec["v"].append(v), rec["photon"].append(ph)
            if q_trap[0] > 0 or q_trap[1] > 0:
                for typ, (tr, ul) in enumerate(source.traps(g, gen, pix, q_trap)):
                    traps.append((sl[tr], pix[tr], t[tr] - life[typ] * F.log(ul), ph[tr], g[tr],
                                  np.full(int(tr.sum()), gen), np.full(int(tr.sum()), typ)))
            x, y = pix // NPIX, pix % NPIX
            nx = np.stack([x - 1, x + 1, x, x], 1)
            ny = np.stack([y, y, y - 1, y + 1], 1)
            ok = np.stack([x > 0, x + 1 < cfg.n_pixels_x, y > 0, y + 1 < cfg.n_pixels_y], 1)
            ok &= source.xtalk(g, gen, pix) < q_ct
            r = np.nonzero(ok)
            cs, cp, ct, cph, cg = sl[r[0]], (nx * NPIX + ny)[r], t[r[0]], ph[r[0]], g[r[0]]
            if len(cs):
                key = cs * (NPIX * NPIX) + cp
                _, first, mult = np.unique(key, return_index=True, return_counts=True)
                front = (cs[first], cp[first], ct[first], cph[first], mult, cg[first])
            else:
                front = (cs, cp, ct, cph, np.zeros(0, np.int64), cg)
            gen += 1
            if gen > MAX_GEN and len(front[0]):
                raise RuntimeError("SiPM cross-talk cascade deeper than the draw layout allows")
    if traps:
        ts, tp, tt, tph, tg, tgen, ttyp = (np.concatenate([x[i] for x in traps]) for i in range(7))
        for i in np.lexsort((tt, ts)):
            sl, pp, t1 = ts[i], tp[i], tt[i]
            if t1 > t_end:
                continue
            v = cfg.overvoltage * (1.0 - F.exp(-(t1 - last[sl, pp]) / cfg.time_constant))
            if source.after(tg[i], tgen[i], pp, ttyp[i]) < p_fire(cfg, v):
                last[sl, pp] = t1
                rec["slot"].append(np.array([sl])), rec["time"].append(np.array([t1]))
                rec["v"].append(np.array([v])), rec["photon"].append(np.array([tph[i]]))
    slot = np.concatenate(rec["slot"])
    v = np.concatenate(rec["v"])
    out = {"chan": chans[slot], "time": np.concatenate(rec["time"]), "charge": cfg.capacitance * v,
           "pe": v / cfg.overvoltage, "photon": np.concatenate(rec["photon"])}
    o = np.lexsort((out["time"], out["chan"]))
    return {k: a[o] for k, a in out.items()}


def charges(steps, status, dev, opt, sipm_cfg, pmap, seed, t_end=1e5, chan_mult=CHAN_MULT):
    """Steps -> SiPM charges (reference chain.py:151-213 numpy branch, closed-form optics). Returns Q dict."""
    P = photons(steps, status, dev, opt, seed)
    if len(P["chan"]) and int(P["chan"].max()) >= chan_mult:
        raise ValueError("channel exceeds CHAN_MULT")
    label = P["evkey"] * chan_mult + P["chan"]
    src = PhiloxSource(seed, (P["evkey"], P["chan"], P["j"], P["rank"]))
    Q = sipm_simulate(label, P["time"], P["row"], pmap, sipm_cfg, src, t_end=t_end)
    Q["n_photons"] = len(P["time"])
    Q["capped"] = P["capped"]
    return Q
