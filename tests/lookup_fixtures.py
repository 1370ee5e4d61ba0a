"""Synthetic v6 lookup tables, a small CRV geometry and CrvStep tables for the lookup-optics tests (test helper).

The released tables are Mu2e-internal, so the tests build their own: a table of random probabilities and delay
weights on coarse bins, and a scintillation table with a known signal speed and attenuation length (probability
~ exp(-d / L) and the delay peaked at d / v, d the distance to the SiPM end), the construction the reference's own
propagation tests use (crv/tests/test_propagation.py:16-36).
"""
import numpy as np

from crvhip import lookup as LK
from crvhip.detector import CRVGeometry

# counter cross-section of the stock tables' header (reference crv/tests/test_lookupgen.py:152-156)
CROSS = dict(half_thickness=9.64, half_width=25.42, fiber_separation=26.0, hole_radius_x=2.0, hole_radius_y=1.0,
             fiber_radius=0.7, corner_radius=1.75)


def random_table(rng, name="LookupTable_rand", half_length=1000.0, reflector=0, pmax=0.01, ck=200.0, dmax=24):
    """Random probabilities in [0, pmax) and delay weights on coarse bins covering the whole counter."""
    edges = {"x": np.linspace(0.0, CROSS["half_thickness"], 5), "y": np.linspace(-CROSS["half_width"],
                                                                                 CROSS["half_width"], 9),
             "z": np.linspace(-half_length, half_length, 11), "beta": np.linspace(0.4, 1.0, 4),
             "theta": np.linspace(0.0, np.pi, 5), "phi": np.linspace(0.0, np.pi, 4), "r": np.linspace(0.0, 0.7, 4)}
    nb = LK.n_bins(edges)

    def tab(n):
        d = rng.integers(0, 40, (n, dmax)).astype(np.uint8)
        d[rng.random((n, dmax)) < 0.3] = 0
        e = rng.integers(0, 40, (n, 3)).astype(np.uint8)
        return LK.Table(rng.uniform(0.0, pmax, n).astype(np.float32), d, d.sum(1, dtype=np.int32), e,
                        e.sum(1, dtype=np.int32))

    # scintillation probabilities fall with the distance to the SiPM end (z fastest in a bin index), so the
    # table's attenuation length can be measured
    S = tab(nb["scint"])
    zc = 0.5 * (edges["z"][1:] + edges["z"][:-1]) + half_length
    S.prob = (S.prob * np.tile(np.exp(-zc / 3000.0), nb["scint"] // len(zc))).astype(np.float32)
    cmap = np.array([[0.5, 0.0], [0.8, 0.6 * ck], [1.0, ck]])
    return LK.LookupTable(name, (6, 0), reflector, CROSS["half_thickness"], CROSS["half_width"], half_length,
                          CROSS["fiber_separation"], CROSS["hole_radius_x"], CROSS["hole_radius_y"],
                          CROSS["fiber_radius"], CROSS["corner_radius"], 0.126, 0.0, cmap, cmap.copy(), edges,
                          {"scint": S, "ckscint": tab(nb["ckscint"]), "ckfiber": tab(nb["ckfiber"])})


def propagation_table(inv_v=1 / 160.0, att=6000.0, half=1000.0, name="LookupTable_prop", reflector=0):
    """Scintillation bins with prob = 0.01 exp(-d / att) and all delay weight at round(d inv_v) + 2, d = z + half;
    Cerenkov tables of the right size that never give photons (zero Cerenkov maps)."""
    edges = {"x": np.linspace(0, 10, 3), "y": np.linspace(-26, 26, 3), "z": np.linspace(-half, half, 41),
             "beta": np.linspace(0.4, 1, 3), "theta": np.linspace(0, np.pi, 3), "phi": np.linspace(0, np.pi, 3),
             "r": np.linspace(0, 0.7, 3)}
    d = 0.5 * (edges["z"][1:] + edges["z"][:-1]) + half
    nz, D = len(d), 40
    prob = np.tile(0.01 * np.exp(-d / att), 4).astype(np.float32)
    delays = np.zeros((4 * nz, D), np.uint8)
    delays[np.arange(4 * nz), np.tile(np.rint(d * inv_v).astype(int) + 2, 4)] = 200
    emis = np.ones((4 * nz, 1), np.uint8)
    S = LK.Table(prob, delays, delays.sum(1, dtype=np.int32), emis, emis.sum(1, dtype=np.int32))
    nb = LK.n_bins(edges)

    def empty(n):
        return LK.Table(np.zeros(n, np.float32), np.ones((n, 1), np.uint8), np.ones(n, np.int32),
                        np.ones((n, 1), np.uint8), np.ones(n, np.int32))

    zero = np.array([[0.5, 0.0], [1.0, 0.0]])
    return LK.LookupTable(name, (6, 0), reflector, 9.64, 25.42, half, 26.0, 2.0, 1.0, 0.7, 1.75, 0.126, 0.0, zero,
                          zero.copy(), edges, {"scint": S, "ckscint": empty(nb["ckscint"]),
                                               "ckfiber": empty(nb["ckfiber"])})


def geometry(half_length=1000.0, sectors=(("EX", 8), ("T1", 8), ("M1", 8))):
    """Bars of the given sectors (name, bars; two modules each), counters along world z, width along x, thickness
    along y, like the stock CRV-top modules; centres spread in x."""
    names = [n for n, _ in sectors]
    sector = np.concatenate([np.full(k, i) for i, (_, k) in enumerate(sectors)]).astype(np.int64)
    nb = len(sector)
    module = np.concatenate([np.arange(k) * 2 // k for _, k in sectors]).astype(np.int64)
    center = np.column_stack([np.arange(nb) * 60.0, np.zeros(nb), np.zeros(nb)])
    halflen = np.tile([2 * CROSS["half_width"] / 2, CROSS["half_thickness"], half_length], (nb, 1))
    return CRVGeometry(names, sector, module, np.zeros(nb, np.int64), np.arange(nb, dtype=np.int64), center, halflen,
                       np.tile([1, 0, 2], (nb, 1)))


def steps(rng, geom, n, n_events=5, short=0.1, nan=0.0, half_length=1000.0, fibre=0.03):
    """A CrvStep table: straight segments inside random bars (muon-like, some protons and electrons), a fraction
    `short` of 0.1 mm steps with large deposits (many arrivals per substep: the chunked Binomial draws), a fraction
    `fibre` running along a fibre's axis (fibre Cerenkov light) and a fraction `nan` with a NaN end time. Returns
    (dict of columns, pdg)."""
    nb = len(geom.sector)
    bar = rng.integers(0, nb, n)
    c = geom.center[bar]
    # local (thickness, width, length) -> world (y, x, z)
    a = np.column_stack([rng.uniform(-9, 9, n), rng.uniform(-25, 25, n), rng.uniform(-0.95, 0.95, n) * half_length])
    length = rng.uniform(1.0, 30.0, n)
    sh = rng.random(n) < short
    length[sh] = 0.1
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    fb = (rng.random(n) < fibre) & ~sh
    a[fb, 0] = rng.uniform(-0.3, 0.3, fb.sum())
    a[fb, 1] = np.where(rng.random(fb.sum()) < 0.5, -13.0, 13.0) + rng.uniform(-0.3, 0.3, fb.sum())
    u[fb] = [0.0, 0.0, 1.0]
    b = a + u * length[:, None]
    b[:, 0] = np.clip(b[:, 0], -9.6, 9.6)
    b[:, 1] = np.clip(b[:, 1], -25.4, 25.4)
    w0 = c + a[:, [1, 0, 2]]
    w1 = c + b[:, [1, 0, 2]]
    edep = rng.uniform(0.05, 3.0, n)
    edep[sh] = rng.uniform(2.0, 6.0, sh.sum())
    t0 = rng.uniform(0, 1500, n)
    t1 = t0 + length * 0.0035
    t1[rng.random(n) < nan] = np.nan
    p = rng.uniform(50, 2000, n)
    pdg = rng.choice([13, -13, 2212, 11, 22], n, p=[0.45, 0.35, 0.1, 0.05, 0.05])
    ev = np.sort(rng.integers(0, n_events, n))
    cols = {"run": np.ones(n, np.int64), "subrun": np.zeros(n, np.int64), "event": ev, "bar": bar, "edep": edep,
            "t0": t0, "t1": t1, "x0": w0[:, 0], "y0": w0[:, 1], "z0": w0[:, 2], "x1": w1[:, 0], "y1": w1[:, 1],
            "z1": w1[:, 2], "px0": p * u[:, 1], "py0": p * u[:, 0], "pz0": p * u[:, 2], "p1": p * 0.999,
            "len": np.linalg.norm(b - a, axis=1), "sp": np.arange(n)}
    return cols, pdg
