"""The synthetic CRV detector: counters, readout channels, calibration conditions and the finder's sector geometry.

5,504 counters of 5.1 m x 5.0 cm x 2.0 cm scintillator, 86 modules x 4 layers x 16 counters, two fibres per
counter read at both ends (4 SiPMs: channel = 4 * counter + sipm, readout side = sipm % 2, fibre = sipm // 2).
Counters run along z; in a module the 16 counters of a layer sit side by side in x (50 mm pitch), the layers are
stacked in y (20 mm pitch); modules are 850 mm apart in x. Conditions are the reference's nominal calibration
(crvfastsim/runconfig.py:229-240: pedestal 100 ADC, pulse area 394.6 ADC ns, pulse height 11.4 ADC, no time
offset, status 0) with 64-channel FEBs, 24 FEBs per ROC. The coincidence finder sees one sector ("CRV") of one
sector type, the structure of the reference's SectorGeom (reco/coinc.py:64-79).
"""
from dataclasses import dataclass

import numpy as np

N_MODULES = 86
N_LAYERS = 4
N_COLS = 16
N_COUNTERS = N_MODULES * N_LAYERS * N_COLS
N_CHANNELS = 4 * N_COUNTERS
COUNTER_WIDTH = 50.0
LAYER_PITCH = 20.0
MODULE_PITCH = N_COLS * COUNTER_WIDTH + 50.0
HALF_LENGTH = 2550.0
CHANNELS_PER_FEB = 64
FEBS_PER_ROC = 24


def counter_id(module, layer, col):
    return (np.asarray(module) * N_LAYERS + np.asarray(layer)) * N_COLS + np.asarray(col)


def decompose(counter):
    counter = np.asarray(counter)
    return counter // (N_LAYERS * N_COLS), (counter // N_COLS) % N_LAYERS, counter % N_COLS


@dataclass
class Conditions:
    """Per-channel CRVCalib constants and channel map (reference chain.py:29-52)."""
    pedestal: np.ndarray
    pulsearea: np.ndarray
    pulseheight: np.ndarray
    timeoffset: np.ndarray
    status: np.ndarray
    roc: np.ndarray
    feb: np.ndarray
    deviation: np.ndarray
    readout_mask: np.ndarray | None = None


def nominal_conditions(n_channels=N_CHANNELS):
    ch = np.arange(n_channels)
    z = np.zeros(n_channels)
    febg = ch // CHANNELS_PER_FEB
    return Conditions(z + 100.0, z + 394.6, z + 11.4, z.copy(), np.zeros(n_channels, np.int64),
                      (febg // FEBS_PER_ROC).astype(np.int64), (febg % FEBS_PER_ROC).astype(np.int64), z.copy())


def effective_status(cond):
    """Reference readout.py:58-72: OR bit 0 into the status of channels the readout mask marks absent."""
    if cond.readout_mask is None:
        return cond.status
    st = np.asarray(cond.status).copy()
    st[~np.asarray(cond.readout_mask, bool)] |= 1
    return st


@dataclass
class SectorGeom:
    """Per-bar lookups of the coincidence finder (reference reco/coinc.py:64-79)."""
    sector: np.ndarray
    layer: np.ndarray
    counter_in_type: np.ndarray
    x: np.ndarray
    y: np.ndarray
    pos: np.ndarray
    length_dir: np.ndarray
    sector_type: np.ndarray
    sector_names: list
    pos_side0: np.ndarray
    pos_side1: np.ndarray
    sipms_side0: np.ndarray
    sipms_side1: np.ndarray


def sector_geom():
    c = np.arange(N_COUNTERS)
    module, layer, col = decompose(c)
    x = module * MODULE_PITCH + col * COUNTER_WIDTH + 0.5 * COUNTER_WIDTH
    y = layer * LAYER_PITCH + 0.5 * LAYER_PITCH
    return SectorGeom(np.zeros(N_COUNTERS, np.int64), layer.astype(np.int64), (module * N_COLS + col).astype(np.int64),
                      x.astype(float), y.astype(float), np.column_stack([x, y, np.zeros(N_COUNTERS)]).astype(float),
                      np.array([2]), np.array([1]), ["CRV"], np.array([-HALF_LENGTH]), np.array([HALF_LENGTH]),
                      np.array([True]), np.array([True]))


CRV_GEOMETRY_FIELDS = ("sector_names", "sector", "module", "layer", "counter", "center", "halflen", "local_to_world")


def is_crv_geometry(g):
    """A reference CRVGeometry (crvfastsim/geometry.py:101-118) or any object with its fields."""
    return g is not None and all(hasattr(g, f) for f in CRV_GEOMETRY_FIELDS)


def sector_geom_from(geom, counters_cfg):
    """The finder's per-bar lookups of a CRVGeometry and its parsed counters file (reference reco/coinc.py:82-110):
    a bar's counter number within its sector type counts the bars of the sectors that precede its sector
    (crs.precedingSectorFor<name> chains), x / y are its centre along the width / thickness direction, and the
    readout ends of a sector lie at its first bar's centre -/+ half its length."""
    c = counters_cfg
    names = list(geom.sector_names)
    need = [f"crs.{k}{n}" for n in names for k in ("sectorType", "precedingSectorFor", "nCountersPerModule",
                                                   "nModules")]
    missing = [k for k in need if k not in c]
    if missing:
        raise ValueError(f"counters_cfg lacks {missing[:4]}{' ...' if len(missing) > 4 else ''}")
    stype = np.array([int(c[f"crs.sectorType{n}"]) for n in names], np.int64)
    prev = np.array([int(c[f"crs.precedingSectorFor{n}"]) for n in names], np.int64)
    per_mod = np.array([int(c[f"crs.nCountersPerModule{n}"]) for n in names], np.int64)
    n_mod = np.array([int(c[f"crs.nModules{n}"]) for n in names], np.int64)
    before = np.zeros(len(names), np.int64)
    for s in range(len(names)):
        p, seen = prev[s], 0
        while p != -1:
            before[s] += n_mod[p] * per_mod[p]
            p = prev[p]
            seen += 1
            if seen > len(names):
                raise ValueError("counters_cfg: crs.precedingSectorFor chains form a loop")
    sector = np.asarray(geom.sector, np.int64)
    l2w = np.asarray(geom.local_to_world, np.int64)
    center = np.asarray(geom.center, float)
    bars = np.arange(len(sector))
    x = center[bars, l2w[:, 1]]
    y = center[bars, l2w[:, 0]]
    first = np.array([np.flatnonzero(sector == s)[0] for s in range(len(names))], np.int64)
    ldir = l2w[first, 2]
    half = np.asarray(geom.halflen, float)[first, ldir]
    mid = center[first, ldir]
    side0 = np.array([bool(c.get(f"crs.sipmsAtSide0{n}", True)) for n in names])
    side1 = np.array([bool(c.get(f"crs.sipmsAtSide1{n}", True)) for n in names])
    cit = before[sector] + per_mod[sector] * np.asarray(geom.module, np.int64) + np.asarray(geom.counter, np.int64)
    return SectorGeom(sector, np.asarray(geom.layer, np.int64), cit, x, y, center, ldir, stype, names, mid - half,
                      mid + half, side0, side1)


def step_longitudinal(geom, bar, pos0, pos1):
    """Midpoint of a step along its counter's length, from the counter centre (mm): world positions [S, 3] of the
    step's ends and the counter's length axis (local_to_world[:, 2]), as the reference's step_inputs places a step
    in its bar (response/photons.py:292-302)."""
    bar = np.asarray(bar, np.int64)
    ldir = np.asarray(geom.local_to_world, np.int64)[bar, 2]
    rows = np.arange(len(bar))
    mid = 0.5 * (np.asarray(pos0, float)[rows, ldir] + np.asarray(pos1, float)[rows, ldir])
    return mid - np.asarray(geom.center, float)[bar, ldir]


def bar_half_lengths(geom):
    """Half length of every counter along its length axis (mm)."""
    ldir = np.asarray(geom.local_to_world, np.int64)[:, 2]
    return np.asarray(geom.halflen, float)[np.arange(len(ldir)), ldir]


@dataclass
class CRVGeometry:
    """The fields of the reference's CRVGeometry (crvfastsim/geometry.py:101-118): per bar its sector, module, layer,
    counter, centre and half lengths (mm, world axes) and the world axes of its (thickness, width, length)."""
    sector_names: list
    sector: np.ndarray
    module: np.ndarray
    layer: np.ndarray
    counter: np.ndarray
    center: np.ndarray
    halflen: np.ndarray
    local_to_world: np.ndarray

    @property
    def nbars(self):
        return len(self.sector)


def synthetic_crv_geometry():
    """The synthetic detector as a CRVGeometry and its counters-file settings: counters along z, side by side in x,
    layers stacked in y; one sector "CRV" of type 1 with 86 modules of 16 counters per layer."""
    c = np.arange(N_COUNTERS)
    module, layer, col = decompose(c)
    G = sector_geom()
    halflen = np.tile([0.5 * COUNTER_WIDTH, 0.5 * LAYER_PITCH, HALF_LENGTH], (N_COUNTERS, 1))
    geom = CRVGeometry(["CRV"], np.zeros(N_COUNTERS, np.int64), module.astype(np.int64), layer.astype(np.int64),
                       col.astype(np.int64), G.pos.copy(), halflen, np.tile([1, 0, 2], (N_COUNTERS, 1)))
    counters = {"crs.sectorTypeCRV": 1, "crs.precedingSectorForCRV": -1, "crs.nCountersPerModuleCRV": N_COLS,
                "crs.nModulesCRV": N_MODULES, "crs.sipmsAtSide0CRV": True, "crs.sipmsAtSide1CRV": True}
    return geom, counters
