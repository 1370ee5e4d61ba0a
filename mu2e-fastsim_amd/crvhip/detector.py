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
