"""Lookup-table optics without a GPU: the v6 reader and writer, the reference's photons.py helper vectors, the
Binomial sampler and the oracle's substep loop against what the mathematics fixes.

Pins (reference tests re-expressed): find_bin (crv/tests/test_photons.py:13-16), Cerenkov interpolation (:19-22),
particles and beta (:25-29), header layout and empty bin records (crv/tests/test_lookupgen.py:118-156), table
propagation (crv/tests/test_propagation.py:39-41), bar keys and ratios (:57-63), the attenuation-ratio slope and the
unchanged centre yield (:91-103), the speed-ratio delay (:123-145). The reference's photons.py does not import here
(its lookup module needs hepsimlib), so the substep loop's parity with it is statistical: the Binomial law (chi2
against the pmf), Binomial thinning of the photon numbers, the delay distribution of a bin, the scintillator and
fibre cross-sections (areas against closed forms).
"""
import struct

import numpy as np
import pytest
from scipy import stats

from crvhip import lookup as LK, photons as PH
from oracle import fmath as F, lookup_optics as LO, philox as PX, rand as R

import lookup_fixtures as FX


# ---------------------------------------------------------------------------------------------------------------------
# v6 file format

def test_v6_round_trip_and_header_layout(tmp_path):
    T = FX.random_table(np.random.default_rng(1), reflector=1, half_length=1185.0)
    path = tmp_path / "LookupTable_rt"
    LK.write(T, path)
    b = path.read_bytes()
    assert struct.unpack_from("<3i", b) == (6, 0, 1)
    assert struct.unpack_from("<10d", b, 16)[:8] == pytest.approx((9.64, 25.42, 1185.0, 26.0, 2.0, 1.0, 0.7, 1.75))
    assert b[12:16] == b"\0\0\0\0"                               # struct padding before the doubles
    U = LK.load("rt", str(tmp_path))
    assert U.version == (6, 0) and U.reflector == 1 and U.half_length == 1185.0
    for k in LK.EDGES:
        assert np.array_equal(U.edges[k], T.edges[k])
    assert np.array_equal(U.ck_scint, T.ck_scint) and np.array_equal(U.ck_fiber, T.ck_fiber)
    for t in LK.TABLES:
        a, c = T.tables[t], U.tables[t]
        assert np.array_equal(a.prob, c.prob) and np.array_equal(a.delay_sum, c.delay_sum)
        m = min(a.delays.shape[1], c.delays.shape[1])
        assert np.array_equal(a.delays[:, :m], c.delays[:, :m]) and not c.delays[:, m:].any()
        assert np.array_equal(a.emis_sum, c.emis_sum)
    assert U.provenance["size"] == len(b) and U.provenance["status"] == "unpinned"


def test_empty_bin_record_and_corrupt_files():
    T = FX.propagation_table()
    T.tables["ckscint"].delays[:] = 0
    T.tables["ckscint"].emis[:] = 0
    b = LK.to_bytes(T)
    rec = struct.pack("<IfQQ", 0, 0.0, 0, 0)                        # a bin without arrivals (reference writer)
    assert rec in b
    with pytest.raises(ValueError, match="trailing"):
        LK.from_bytes(b + b"\0")
    bad = bytearray(b)
    bad[0:4] = struct.pack("<i", 5)
    with pytest.raises(ValueError, match="version"):
        LK.from_bytes(bytes(bad))
    with pytest.raises(ValueError):
        LK.from_bytes(b[:200])


# ---------------------------------------------------------------------------------------------------------------------
# reference helper vectors

def test_find_bin_matches_stock_inclusive_edges():
    e = np.array([0.0, 1.0, 2.0, 3.0])
    v = np.array([-0.1, 0.0, 0.5, 1.0, 2.999, 3.0, 3.1, np.nan])
    want = [-1, 0, 0, 0, 2, 2, -1, -1]
    assert PH.find_bin(e, v).tolist() == want and LO.find_bin(e, v).tolist() == want


def test_cerenkov_interp_and_charge():
    cmap = np.array([[0.5, 0.0], [0.6, 10.0], [1.0, 50.0]])
    beta, q = np.array([0.25, 0.55, 0.8, 1.2]), np.array([1, 1, -2, 1])
    assert np.allclose(PH.cerenkov_per_mm(beta, q, cmap), [0.0, 5.0, 60.0, 50.0])
    assert np.allclose(LO.interp_map(beta, cmap) * np.abs(q), [0.0, 5.0, 60.0, 50.0])
    b = np.random.default_rng(2).uniform(-0.2, 1.3, 1000)
    assert np.allclose(LO.interp_map(b, cmap), PH.cerenkov_per_mm(b, 1.0, cmap), rtol=1e-14, atol=1e-13)


def test_beta_and_particles():
    m, q, known = PH.particle_props(np.array([13, -11, 22, 2212, 1000020040, 999]))
    assert q.tolist() == [-1, 1, 0, 1, 2, 0] and known.tolist() == [True] * 5 + [False]
    assert m[4] == pytest.approx(4 * PH.AMU)
    b = PH.beta_of(np.array([105.66, 0.0]), np.array([105.66, 0.0]), np.array([105.66, 0.0]))
    assert np.isclose(b[0], 1 / np.sqrt(2)) and b[1] == 1.0
    assert np.array_equal(LO.beta_of(np.array([105.66, 3.0]), np.array([105.66, 2.0]), np.array([105.66, 0.0])),
                          PH.beta_of([105.66, 3.0], [105.66, 2.0], [105.66, 0.0]))


def test_table_propagation_recovers_known_values():
    T = FX.propagation_table()
    inv_v, att = PH.table_propagation(T)
    assert abs(1 / inv_v - 160.0) < 2.0 and abs(att - 6000.0) < 30.0
    assert LO.table_propagation(T) == (inv_v, att)                  # the oracle's fit gives the same bits


class _G:
    sector_names = ["EX", "T1", "T2"] + [f"M{i}" for i in range(1, 9)]
    sector = np.r_[np.zeros(256), np.ones(128), np.full(128, 2), np.repeat(np.arange(3, 11), 8)].astype(int)
    module = np.r_[np.arange(256) // 64, np.arange(128) // 64, np.arange(128) // 64, np.zeros(64)].astype(int)


def test_bar_keys_and_ratios():
    sk, mk = PH.bar_keys(_G)
    assert mk[0] == "EX0" and mk[255] == "EX3" and mk[256] == "T10" and mk[511] == "T21" and mk[512] == "M1"
    r = PH.bar_ratios({"EX": 1.08, "EX3": 1.1, "TAG": 0.9, "M2": 0.8}, _G)
    assert r[0] == 1.08 and r[200] == 1.1 and r[300] == 1.0 and r[512] == 0.9 and r[520] == 0.8
    with pytest.raises(KeyError):
        PH.bar_ratios({"EX7": 1.0}, _G)
    with pytest.raises(ValueError):
        PH.bar_ratios({"EX": 0.0}, _G)


def test_bar_arrays_fail_closed():
    from crvhip.config import ResponseConfig, SectorConfig
    G = FX.geometry()
    T = {"LookupTable_rand": FX.random_table(np.random.default_rng(0))}
    ok = ResponseConfig(sectors={n: SectorConfig("rand", 0, 1000.0) for n in G.sector_names})
    bars, tl = PH.bar_arrays(G, ok, T)
    assert (bars["table"] == 0).all() and len(tl) == 1
    with pytest.raises(ValueError, match="no entry"):
        PH.bar_arrays(G, ResponseConfig(sectors={"EX": SectorConfig("rand", 0, 1.0)}), T)
    with pytest.raises(ValueError, match="not among"):
        PH.bar_arrays(G, ResponseConfig(sectors={n: SectorConfig("nope", 0, 1.0) for n in G.sector_names}), T)
    with pytest.raises(ValueError, match="reflector"):
        PH.bar_arrays(G, ResponseConfig(sectors={n: SectorConfig("rand", 1, 1.0) for n in G.sector_names}), T)


# ---------------------------------------------------------------------------------------------------------------------
# samplers and math

def test_acos_within_two_ulp():
    x = np.r_[np.linspace(-1, 1, 100001), -1.0, 1.0, 0.5, -0.5]
    r, ref = F.acos(x), np.arccos(x)
    assert (np.abs(r - ref) <= 2 * np.spacing(np.abs(ref))).all()
    assert F.acos(np.array([-1.0]))[0] == np.pi and F.acos(np.array([1.0]))[0] == 0.0
    assert np.isnan(F.acos(np.array([1.5, np.nan]))).all()


@pytest.mark.parametrize("n,p", [(1, 0.3), (7, 0.5), (40, 0.02), (300, 0.05), (5000, 0.01), (60, 0.93),
                                 (200000, 0.0004), (30, 1e-9)])
def test_binomial_law(n, p):
    N = 60000
    rng = np.random.default_rng(n + int(p * 1e6))
    k, ok = R.binomial(np.full(N, n), np.full(N, p), lambda c: rng.random(N))
    assert ok.all() and k.min() >= 0 and k.max() <= n
    mean, var = n * p, n * p * (1 - p)
    assert abs(k.mean() - mean) < 5 * np.sqrt(var / N) + 1e-12
    if var > 0.5:
        assert abs(k.var() / var - 1) < 0.05
    lo, hi = int(stats.binom.ppf(1e-4, n, p)), int(stats.binom.ppf(1 - 1e-4, n, p))
    edges = np.arange(lo, hi + 2)
    obs = np.histogram(k, edges)[0]
    exp = N * stats.binom.pmf(edges[:-1], n, p)
    m = exp > 10
    chi2 = ((obs[m] - exp[m]) ** 2 / exp[m]).sum()
    assert stats.chi2.sf(chi2, max(m.sum() - 1, 1)) > 1e-4


def test_binomial_edges():
    u = lambda c: np.full(5, 0.5)  # noqa: E731
    k, ok = R.binomial(np.array([0, 10, 10, 10, 3]), np.array([0.5, 0.0, 1.0, np.nan, 1.0]), u)
    assert k.tolist() == [0, 0, 10, 0, 3] and ok.all()
    # chunking: n q > 16 needs several chunks, each with its own uniform
    calls = []
    R.binomial(np.array([10000]), np.array([0.01]), lambda c: calls.append(c) or np.array([0.5]))
    assert calls == list(range(len(calls))) and len(calls) >= 6


# ---------------------------------------------------------------------------------------------------------------------
# the oracle's substep loop against closed forms

def _bars(G, tid=0, refl=0, Y=1000.0, speed=1.0, atten=1.0):
    nb = len(G.sector)
    return {"table": np.full(nb, tid), "reflector": np.full(nb, refl), "yield": np.full(nb, Y),
            "speed_ratio": np.full(nb, speed), "atten_ratio": np.full(nb, atten), "center": G.center,
            "l2w": G.local_to_world}


def _line_steps(G, T, n, rng, y=12.0):
    """Point-like steps (0.05 mm along the thickness) at uniform z in bar 0, 200 scintillation photons each."""
    z = rng.uniform(-0.95, 0.95, n) * T.half_length
    p0 = np.column_stack([np.full(n, y), rng.uniform(-9, 9, n) * 0 + 1.0, z])  # world (x=width, y=thickness, z)
    p1 = p0.copy()
    p1[:, 1] += 0.05
    return {"evkey": np.arange(n), "bar": np.zeros(n, np.int64), "t0": np.zeros(n), "t1": np.zeros(n),
            "pos0": p0, "pos1": p1, "edep": np.full(n, 0.2), "len": np.full(n, 0.05), "p0": np.full(n, 1e4),
            "p1": np.full(n, 1e4), "mass": np.full(n, 105.66), "charge": np.full(n, -1.0)}, z


def _slope(z, zsel, half):
    zb = np.linspace(-half, half, 11)
    n = np.histogram(zsel, zb)[0] / np.histogram(z, zb)[0]
    return np.polyfit(0.5 * (zb[1:] + zb[:-1]), np.log(n), 1)[0]


def test_oracle_attenuation_ratio_shifts_the_log_slope():
    G = FX.geometry()
    T = FX.propagation_table()
    props = [LO.table_propagation(T)]
    st, z = _line_steps(G, T, 40000, np.random.default_rng(101))
    status, dev = np.zeros(4 * len(G.sector), np.int64), np.ones(4 * len(G.sector))
    base = LO.photons(st, _bars(G), [T], props, np.zeros(len(dev)), status, 1.0, 1.0, 7)
    r = 0.6
    att = LO.photons(st, _bars(G, atten=r), [T], props, np.zeros(len(dev)), status, 1.0, 1.0, 7)
    mu = (1 / r - 1) / props[0][1]
    s0 = base["chan"] == 0
    s1 = att["chan"] == 0
    # SiPM 0 reads the end at local z = -H: the table frame is z; extra log slope -mu, centre yield unchanged
    d = _slope(z, z[base["row"][s0]], T.half_length) - _slope(z, z[att["row"][s1]], T.half_length)
    assert abs(d - mu) < 0.25 * mu
    zc = np.abs(z) < 0.05 * T.half_length
    n0, n1 = zc[base["row"][s0]].sum(), zc[att["row"][s1]].sum()
    assert abs(n1 / n0 - 1) < 5 / np.sqrt(n0)


def test_oracle_speed_ratio_delays_both_ends():
    G = FX.geometry()
    T = FX.propagation_table()
    props = [LO.table_propagation(T)]
    st, z = _line_steps(G, T, 300, np.random.default_rng(7))
    nch = 4 * len(G.sector)
    for r in (1.0, 1.25):
        Rw = LO.rows(st, _bars(G, speed=r), [T], props, np.zeros(nch), np.zeros(nch, np.int64), 1.0, 1.0, 3)
        s0 = Rw["s"] == 0
        k = props[0][0] * (1 / r - 1)
        assert np.allclose(Rw["ta"][s0], k * (Rw["a"][s0, 2] + T.half_length), rtol=0, atol=1e-9)
        assert abs(np.polyfit(Rw["a"][s0, 2], Rw["ta"][s0], 1)[0] - k) < 1e-12


def test_oracle_arrivals_thin_the_photon_numbers_and_draw_bin_delays():
    """Every arrival comes from a substep with a valid bin; per type the arrival count is Binomial(n, p) summed over
    substeps (mean sum n p, checked at 5 sigma); delays within a bin follow its weights (chi2)."""
    rng = np.random.default_rng(5)
    G = FX.geometry()
    T = FX.random_table(rng)
    props = [(1 / 160.0, 6000.0)]
    cols, pdg = FX.steps(rng, G, 400, short=0.05)
    m, q, _ = PH.particle_props(pdg)
    ev = cols["event"]
    st = {"evkey": ev, "bar": cols["bar"], "t0": cols["t0"], "t1": cols["t1"],
          "pos0": np.column_stack([cols[f"{a}0"] for a in "xyz"]), "pos1": np.column_stack([cols[f"{a}1"] for a in "xyz"]),
          "edep": cols["edep"], "len": cols["len"], "p0": np.sqrt(cols["px0"] ** 2 + cols["py0"] ** 2 + cols["pz0"] ** 2),
          "p1": cols["p1"], "mass": m, "charge": q}
    nch = 4 * len(G.sector)
    P = LO.photons(st, _bars(G), [T], props, np.zeros(nch), np.zeros(nch, np.int64), 1.0, 1.0, 11)
    Rw = P["rows"]
    assert len(P["time"]) > 1000 and (P["bin"] >= 0).all()
    same = (np.diff(P["row"]) == 0) & (np.diff(P["chan"]) == 0)
    assert (np.diff(P["rank"])[same] == 1).all() and (P["rank"][1:][~same] == 0).all()
    # expected arrivals: recompute the substep probabilities from the oracle's own rows
    nsub = Rw["nsub"]
    ri = np.repeat(np.arange(len(nsub)), nsub)
    kk = np.arange(len(ri)) - np.repeat(np.cumsum(nsub) - nsub, nsub)
    frac = (kk + 0.5) / nsub[ri]
    p = Rw["a"][ri] + Rw["d"][ri] * frac[:, None]
    insc = LO.inside_scint(T, p[:, 0], p[:, 1], p[:, 2])
    infb, r, phi = LO.inside_fiber(T, p[:, 0], p[:, 1], p[:, 2], Rw["d"][ri, 0], Rw["d"][ri, 1])
    infb &= ~insc
    bins = LO._bins(T, p[:, 0], p[:, 1], p[:, 2], Rw["beta"][ri], Rw["theta"][ri], r, phi, insc, infb)
    for typ, name in enumerate(LK.TABLES):
        ok = bins[typ] >= 0
        pr = np.minimum(T.tables[name].prob[bins[typ][ok]].astype(float) * Rw["dev"][ri[ok]], 1.0)
        n = Rw["n"][ri[ok], typ]
        mean, var = (n * pr).sum(), (n * pr * (1 - pr)).sum()
        got = (P["typ"] == typ).sum()
        assert abs(got - mean) <= 5 * np.sqrt(var) + 1, (name, got, mean)
    # delay indices of all arrivals against the mixture of their bins' weights
    exp = np.zeros(64)
    obs = np.zeros(64)
    for typ, name in enumerate(LK.TABLES):
        sel = P["typ"] == typ
        if not sel.any():
            continue
        W = T.tables[name].delays[P["bin"][sel]].astype(float)
        exp[:W.shape[1]] += (W / W.sum(1, keepdims=True)).sum(0)
        assert (P["delay"][sel] < W.shape[1]).all()
        assert (W[np.arange(len(W)), P["delay"][sel]] > 0).all()      # never a zero-weight index
        obs[:W.shape[1]] += np.bincount(P["delay"][sel], minlength=W.shape[1])
    m = exp > 5
    chi2 = ((obs[m] - exp[m]) ** 2 / exp[m]).sum()
    assert stats.chi2.sf(chi2, max(m.sum() - 1, 1)) > 1e-4
    # an arrival's time is its substep's centre time plus the delay index
    row_of = {(int(s_), int(c_)): i for i, (s_, c_) in enumerate(zip(Rw["step"], Rw["chan"]))}
    ridx = np.array([row_of[(int(a_), int(c_))] for a_, c_ in zip(P["row"], P["chan"])])
    tcen = Rw["ta"][ridx] + (Rw["tb"][ridx] - Rw["ta"][ridx]) * ((P["sub"] + 0.5) / nsub[ridx])
    assert np.array_equal(P["time"], tcen + P["delay"])


def test_scintillator_and_fibre_cross_sections():
    """Uniform points over the counter cross-section: the scintillator fraction is the box minus the two elliptical
    holes and the four rounded corners; the fibre fraction is two fibre discs (each inside its hole)."""
    T = FX.propagation_table()
    rng = np.random.default_rng(9)
    N = 2_000_000
    x = rng.uniform(-T.half_thickness, T.half_thickness, N)
    y = rng.uniform(-T.half_width, T.half_width, N)
    z = np.zeros(N)
    A = 4 * T.half_thickness * T.half_width
    holes = 2 * np.pi * T.hole_radius_x * T.hole_radius_y
    corners = 4 * (T.corner_radius ** 2 - np.pi * T.corner_radius ** 2 / 4)
    f = LO.inside_scint(T, x, y, z).mean()
    fe = 1 - (holes + corners) / A
    assert abs(f - fe) < 5 * np.sqrt(fe * (1 - fe) / N)
    fb, r, _ = LO.inside_fiber(T, x, y, z, np.ones(N), np.zeros(N))
    fbe = 2 * np.pi * T.fiber_radius ** 2 / A
    assert abs(fb.mean() - fbe) < 5 * np.sqrt(fbe / N)
    assert (r[fb] <= T.fiber_radius).all()


def test_draw_layout_kinds_are_distinct():
    assert len({PX.KIND_LCOUNT, PX.KIND_LBINOM, PX.KIND_LDELAY, PX.KIND_AFTER, PX.KIND_TRAP}) == 5
    assert max(PX.KIND_LCOUNT, PX.KIND_LBINOM, PX.KIND_LDELAY) < 16
