"""A synthetic workload for the lookup-table optics: one v6 table for the synthetic detector and the cosmic steps of
crvhip.synth as CrvStep rows.

The released tables are Mu2e-internal, so the table is made up, on the stock tables' layout (reference
response/lookup.py:79-180): scintillation bins over (x, y, z) of the counter, scintillator-Cerenkov bins over
(x, y, z, beta), fibre-Cerenkov bins over (beta, theta, phi, r, z), each with an arrival probability at SiPM 0 and
uint8 delay and fibre-emission weights. Probabilities fall with the distance to the SiPM end with the two-exponential
attenuation of the closed-form optics and are scaled so that a minimum-ionising step gives about the closed-form
optics' photoelectrons (39 PE/MeV at 1 m); delays peak at the propagation time d / v plus an exponential emission
tail. Sizes are set by the bin counts (the default table holds 405K bins, about 40 MB of delay weights).
"""
import numpy as np

from . import config as CF, detector as DT, lookup as LK

# counter cross-section of the synthetic detector (mm): thickness 20, width 50; fibres 13 mm either side of the centre
HALF_THICKNESS, HALF_WIDTH = 0.5 * DT.LAYER_PITCH, 0.5 * DT.COUNTER_WIDTH
TABLE = "synthetic"


def table(seed=7, nx=8, ny=20, nz=204, nbeta=6, ntheta=6, nphi=6, nr=4, dmax=96, yield_per_mev=39794.0):
    """The synthetic detector's lookup table (reflector 0: SiPMs at both ends)."""
    rng = np.random.default_rng(seed)
    H = DT.HALF_LENGTH
    edges = {"x": np.linspace(0.0, HALF_THICKNESS, nx + 1), "y": np.linspace(-HALF_WIDTH, HALF_WIDTH, ny + 1),
             "z": np.linspace(-H, H, nz + 1), "beta": np.linspace(0.4, 1.0, nbeta + 1),
             "theta": np.linspace(0.0, np.pi, ntheta + 1), "phi": np.linspace(0.0, np.pi, nphi + 1),
             "r": np.linspace(0.0, 0.7, nr + 1)}
    nb = LK.n_bins(edges)
    zc = 0.5 * (edges["z"][1:] + edges["z"][:-1])
    d = zc + H  # distance to SiPM 0's end
    o = CF.OpticsConfig()
    att = o.w_short * np.exp(-d / o.l_short) + o.w_long * np.exp(-d / o.l_long)
    ref = o.w_short * np.exp(-o.ref_distance / o.l_short) + o.w_long * np.exp(-o.ref_distance / o.l_long)
    p_ref = o.pe_per_mev / yield_per_mev  # PE per emitted photon at 1 m

    def weights(n, peak):
        """Delay weights over dmax ns: a peak at `peak` (per bin) with an exponential tail, plus noise."""
        k = np.arange(dmax)[None, :]
        w = np.exp(-np.maximum(k - peak[:, None], 0) / 7.4) * (k >= peak[:, None])
        w = np.rint(200 * w + rng.integers(0, 3, (n, dmax)) * (k >= peak[:, None])).astype(np.uint8)
        w[:, -1] = np.maximum(w[:, -1], 1)  # no empty row
        return w

    def tab(n, prob, peak):
        dl = weights(n, peak)
        e = rng.integers(1, 20, (n, 3)).astype(np.uint8)
        return LK.Table(prob.astype(np.float32), dl, dl.sum(1, dtype=np.int32), e, e.sum(1, dtype=np.int32))

    # scintillation: [x, y, z], z fastest
    per_z = p_ref * att / ref
    s_prob = np.tile(per_z, nb["scint"] // nz) * rng.uniform(0.9, 1.1, nb["scint"])
    s_peak = np.tile(np.minimum(np.rint(d / 160.0).astype(int) + 2, dmax - 20), nb["scint"] // nz)
    S = tab(nb["scint"], s_prob, s_peak)
    # scintillator Cerenkov: [x, y, z, beta], beta fastest
    c_prob = np.repeat(s_prob, nbeta) * 0.05
    C = tab(nb["ckscint"], c_prob, np.repeat(s_peak, nbeta))
    # fibre Cerenkov: [beta, theta, phi, r, z], z fastest
    f_prob = np.tile(0.01 * np.exp(-d / o.l_long), nb["ckfiber"] // nz)
    F = tab(nb["ckfiber"], f_prob, np.tile(s_peak[:nz], nb["ckfiber"] // nz))
    cmap = np.array([[0.6, 0.0], [0.8, 40.0], [1.0, 60.0]])
    return LK.LookupTable(f"LookupTable_{TABLE}", (6, 0), 0, HALF_THICKNESS, HALF_WIDTH, H, 26.0, 2.0, 1.0, 0.7,
                          1.75, 0.126, 0.0, cmap, cmap.copy(), edges, {"scint": S, "ckscint": C, "ckfiber": F})


def response_config():
    """ResponseConfig of the synthetic detector's one sector ("CRV") on the synthetic table."""
    return CF.ResponseConfig(photon_yield_scale=1.0, sectors={"CRV": CF.SectorConfig(TABLE, 0, 39794.0)})


def crv_steps(S, seed=11):
    """The cosmic steps of crvhip.synth (evkey, counter, edep, z, t) as CrvStep rows of vertical muons through the
    synthetic detector's counters (world y across the 20 mm thickness, x uniform across the width); returns (columns,
    pdg)."""
    rng = np.random.default_rng(seed)
    geom, _ = DT.synthetic_crv_geometry()
    bar = np.asarray(S["counter"], np.int64)
    n = len(bar)
    c = geom.center[bar]
    x = c[:, 0] + rng.uniform(-0.9, 0.9, n) * HALF_WIDTH
    z = np.asarray(S["z"], float)
    t = np.asarray(S["t"], float)
    y0, y1 = c[:, 1] - HALF_THICKNESS, c[:, 1] + HALF_THICKNESS
    length = 2 * HALF_THICKNESS
    p = rng.uniform(1000.0, 5000.0, n)  # MeV, vertical
    ev = np.asarray(S["evkey"], np.int64)
    cols = {"run": np.ones(n, np.int64), "subrun": ev >> 20, "event": ev & ((1 << 20) - 1), "bar": bar,
            "edep": np.asarray(S["edep"], float), "t0": t, "t1": t + length / 299.79,
            "x0": x, "y0": y0, "z0": z, "x1": x, "y1": y1, "z1": z,
            "px0": np.zeros(n), "py0": p, "pz0": np.zeros(n), "p1": p - 2.0, "len": np.full(n, length),
            "sp": np.arange(n)}
    pdg = np.where(rng.random(n) < 0.55, 13, -13)
    return cols, pdg
