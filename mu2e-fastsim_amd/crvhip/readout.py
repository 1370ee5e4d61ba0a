"""Physical SiPM presence (reference crvfastsim/readout.py:9-72): channel 4 * bar + sipm exists iff its sector has
SiPMs at side sipm % 2 (counters config crs.sipmsAtSide{0,1}<sector>), independent of calibration status. Absent
sensors get status bit 0, so they take no photons, no dark noise and no pulses."""
from collections.abc import Mapping
from dataclasses import replace

import numpy as np

from .detector import effective_status  # noqa: F401  (the same rule; re-exported for the reference's name)


def _nbars(geom):
    n = getattr(geom, "nbars", None)
    return len(np.asarray(geom.sector)) if n is None else n


def sensor_mask(geom, counters_cfg):
    """Dense bool mask over 4 * nbars channels (reference readout.py:9-50), with the reference's checks."""
    names = getattr(geom, "sector_names", None)
    if (not isinstance(names, (list, tuple, np.ndarray)) or (isinstance(names, np.ndarray) and names.ndim != 1)
            or len(names) == 0):
        raise ValueError("readout geometry requires nonempty sector_names")
    names = list(names)
    if any(not isinstance(n, str) or not n for n in names) or len(set(names)) != len(names):
        raise ValueError("readout geometry sector_names must be unique nonempty strings")
    sector = np.asarray(getattr(geom, "sector", None))
    if sector.ndim != 1 or sector.dtype.kind not in "iu":
        raise ValueError("readout geometry sector must be a one-dimensional integer array")
    nbars = _nbars(geom)
    if (not isinstance(nbars, (int, np.integer)) or isinstance(nbars, (bool, np.bool_)) or nbars < 0
            or sector.shape != (nbars,)):
        raise ValueError("readout geometry sector size must equal nbars")
    if np.any(sector < 0) or np.any(sector >= len(names)):
        raise ValueError("readout geometry contains an unknown sector index")
    if not isinstance(counters_cfg, Mapping):
        raise ValueError("readout counters_cfg must be a mapping")
    if "crs.sectorNames" in counters_cfg:
        configured = counters_cfg["crs.sectorNames"]
        if (not isinstance(configured, (list, tuple, np.ndarray))
                or (isinstance(configured, np.ndarray) and configured.ndim != 1) or list(configured) != names):
            raise ValueError("readout geometry sector_names disagree with crs.sectorNames")
    sides = np.empty((len(names), 2), dtype=bool)
    for i, name in enumerate(names):
        for side in (0, 1):
            key = f"crs.sipmsAtSide{side}{name}"
            if key not in counters_cfg:
                raise ValueError(f"readout counters missing {key}")
            value = counters_cfg[key]
            if not isinstance(value, (bool, np.bool_)):
                raise ValueError(f"readout counters {key} must be boolean, got {value!r}")
            sides[i, side] = value
    return sides[sector][:, [0, 1, 0, 1]].reshape(4 * nbars)


def apply(cond, geom, counters_cfg):
    """Conditions with the physical mask (a copy; channels beyond the bars stay present), reference readout.py:52-56."""
    mask = sensor_mask(geom, counters_cfg)
    n = len(np.asarray(cond.status))
    if n < len(mask):
        raise ValueError(f"readout: conditions cover {n} channels, the geometry has {len(mask)}")
    return replace(cond, readout_mask=np.concatenate([mask, np.ones(n - len(mask), bool)]))
