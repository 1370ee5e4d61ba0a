"""Settings the chain supports, checked before any device work (reference crvfastsim/response/contract.py:5-21).

Everything outside the supported envelope raises ValueError naming the setting, so a caller never gets results of a
configuration other than the one asked for.
"""
import numpy as np

PIXEL_IDS = ("lrint", "truncate")
UNSUPPORTED_SIPM = ("thermal_rate", "trap0_prob", "trap1_prob")


def validate_sipm(sipm):
    """Reference contract.py:5-12: thermal noise and trap afterpulses are refused; pixel ids lrint or truncate."""
    unsupported = [k for k in UNSUPPORTED_SIPM if getattr(sipm, k) != 0]
    if unsupported:
        raise ValueError(f"CRV response: unsupported SiPM processes: {unsupported}")
    if getattr(sipm, "pixel_id", "lrint") not in PIXEL_IDS:
        raise ValueError(f"CRV response: SiPM pixel_id must be one of {PIXEL_IDS}, got {sipm.pixel_id!r}")


def validate_window(window):
    """Reference contract.py:14-21: the response window holds finite, increasing bounds."""
    w = np.asarray(window, dtype=float)
    if w.shape != (2,) or not np.isfinite(w).all() or w[1] <= w[0]:
        raise ValueError("CRV response: window must contain finite, increasing bounds")


def validate_chain(cfg):
    validate_sipm(cfg.sipm)
    validate_window(cfg.window)
    if cfg.backend != "hip":
        raise ValueError(f"ChainConfig.backend must be 'hip', got {cfg.backend!r}")
    if cfg.precision not in ("fp64", "fp32"):
        raise ValueError(f"ChainConfig.precision must be 'fp64' or 'fp32', got {cfg.precision!r}")
