"""The SiPM stage behind the reference's response/sipm.py simulate (sipm.py:102-198), on the HIP kernels of
crv_sipm_run: photon arrivals, dark noise and trap afterpulses -> SiPM charges.

simulate() takes the reference's arguments. The numpy Generator of the reference becomes the seed of the keyed
Philox draws (`key`, or a seed taken from `rng`): a photon's draws are keyed by its label and list position, dark
discharge k of a SiPM by (label, k), so a call is reproducible and independent of launch shape. The arguments are
checked the way the reference's device path checks them (response/sipm_jax.py: labels, windows, rates, pixels) and
every failure raises ValueError naming the argument.
"""
import dataclasses

import numpy as np

from . import _lib as L, config as CF, contract as CT, detector as DT
from .chain import _empty, _host, _seed, _torch

_CTX = {}


def _context(cfg, pmap):
    """A device context for this SiPMConfig (cached by identity), with `pmap` as its photon map."""
    ent = _CTX.get(id(cfg))
    if ent is None:
        _CTX.clear()
        quiet = dataclasses.replace(cfg, thermal_rate=0.0, trap0_prob=0.0, trap1_prob=0.0)  # noise travels per call
        ccfg = dataclasses.replace(CF.ChainConfig(), sipm=quiet)
        from . import template as TP
        ent = [L.Context(ccfg, DT.nominal_conditions(), None, TP.load(ccfg.waveform)), ccfg, cfg]
        _CTX[id(cfg)] = ent
    ctx, ccfg = ent[0], ent[1]
    ctx.use_photon_map(ccfg, pmap)
    return ctx, ccfg


def _ints(name, a, nonneg=True):
    a = np.asarray(a)
    if a.ndim != 1 or (a.size and a.dtype.kind not in "iu"):
        raise ValueError(f"{name} must be a 1-d integer array")
    if a.dtype.kind == "u" and a.size and int(a.max()) > np.iinfo(np.int64).max:
        raise ValueError(f"{name} exceeds the int64 range")
    a = a.astype(np.int64)
    if nonneg and a.size and a.min() < 0:
        raise ValueError(f"{name} must be non-negative")
    return a


def _check_cfg(cfg, t_start, t_end, thermal):
    """The reference device path's checks (response/sipm_jax.py:257-290), plus the trap settings it does not take."""
    for k in ("thermal_rate", "trap0_prob", "trap1_prob"):
        v = getattr(cfg, k)
        if not np.isfinite(v) or v < 0:
            raise ValueError(f"SiPM {k} must be finite and non-negative, got {v!r}")
    for i in (0, 1):
        if getattr(cfg, f"trap{i}_prob") > 0:
            life = getattr(cfg, f"trap{i}_lifetime")
            if not np.isfinite(life) or life <= 0:
                raise ValueError(f"SiPM trap{i}_lifetime must be finite and positive, got {life!r}")
    if cfg.pixel_id not in CT.PIXEL_IDS:
        raise ValueError(f"SiPM pixel_id must be one of {CT.PIXEL_IDS}, got {cfg.pixel_id!r}")
    for k in ("n_pixels_x", "n_pixels_y"):
        v = getattr(cfg, k)
        if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)) or not 1 <= v <= 41:
            raise ValueError(f"SiPM pixel grid {k} must be an integer in 1..41, got {v!r}")
    for x, y in cfg.inactive:
        if not all(isinstance(c, (int, np.integer)) and not isinstance(c, (bool, np.bool_)) and 0 <= c <= 40
                   for c in (x, y)):
            raise ValueError(f"SiPM inactive pixel ({x!r}, {y!r}) outside 0..40")
    if not (np.isfinite(t_start) and np.isfinite(t_end) and 0 <= t_start <= t_end):
        raise ValueError(f"t_start / t_end: require finite 0 <= t_start <= t_end, got {t_start!r}, {t_end!r}")
    pf = cfg.p1 * (1.0 - np.exp(-cfg.overvoltage / cfg.p2))
    if (thermal or cfg.trap0_prob > 0 or cfg.trap1_prob > 0) and not (np.isfinite(pf) and 0 < pf <= 1):
        raise ValueError(f"dark noise and traps need a finite firing probability p_fire(Vov) in (0, 1], got {pf!r}")
    if thermal and not np.isfinite(cfg.thermal_rate / pf * (t_end - t_start)):
        raise ValueError("thermal_rate: the dark-noise Poisson mean must be finite")


def simulate(chan, time, photon, pmap, cfg, rng=None, t_end=1e5, t_start=0.0, thermal_chans=None, chunk=None,
             key=None):
    """SiPM charges for photon arrivals (reference response/sipm.py:102-198).

    chan [P] SiPM label per photon (event key * 32768 + channel), time [P] ns, photon [P] an id carried to the
    charges. thermal_chans: the distinct labels that get dark noise at cfg.thermal_rate over [t_start, t_end)
    (needed when the rate is positive, ignored when it is 0). `chunk` (the reference's SiPMs per numpy pass) does not change results and is not used.
    Returns dict of arrays per charge, sorted by (chan, time): chan, time, charge (C), pe, photon (-1: dark noise).
    """
    label = _ints("chan", chan)
    t = np.asarray(time, np.float64)
    ph = _ints("photon", photon, nonneg=False)
    if t.ndim != 1 or len(t) != len(label) or len(ph) != len(label):
        raise ValueError("chan, time and photon must be 1-d arrays of one length")
    if not np.isfinite(t).all():
        raise ValueError("time must be finite")
    thermal = bool(cfg.thermal_rate > 0)
    tl = np.zeros(0, np.int64)
    if thermal:
        if thermal_chans is None:
            raise ValueError("thermal_chans: a positive thermal_rate needs the labels that get dark noise")
        tl = _ints("thermal_chans", thermal_chans)
        if not len(tl) or len(np.unique(tl)) != len(tl):
            raise ValueError("thermal_chans must be non-empty and distinct")
    _check_cfg(cfg, float(t_start), float(t_end), thermal)
    torch = _torch()
    ctx, ccfg = _context(cfg, pmap)
    ctx.set_params(ccfg, _seed(key, rng), 0)
    dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to("cuda")
    d_label, d_time, d_ph = dev(label, np.int64), dev(t, np.float64), dev(ph, np.int64)
    d_tl = dev(tl, np.int64) if thermal else None
    pl = L.PhotonList(L.ptr(d_label), L.ptr(d_time), L.ptr(d_ph), None, None, len(label))
    nz = L.SipmNoise(float(cfg.thermal_rate) if thermal else 0.0, float(cfg.trap0_prob), float(cfg.trap1_prob),
                     float(cfg.trap0_lifetime), float(cfg.trap1_lifetime), float(t_start), float(t_end),
                     L.ptr(d_tl), len(tl) if thermal else 0)
    cap = 2 * len(label) + 4096
    if thermal:
        pf = cfg.p1 * (1.0 - np.exp(-cfg.overvoltage / cfg.p2))
        cap += int(3 * cfg.thermal_rate / pf * (t_end - t_start) * len(tl)) + 4096
    import ctypes
    for _ in range(2):
        bufs = {k: _empty(cap, dt) for k, dt in (("label", torch.int64), ("time", torch.float64),
                                                 ("charge", torch.float64), ("pe", torch.float64),
                                                 ("photon", torch.int64))}
        out = L.Charges(*[L.ptr(bufs[k]) for k in ("label", "time", "charge", "pe", "photon")], cap, 0, 0, 0)
        rc = ctx.lib.crv_sipm_run(ctx.h, ctypes.byref(pl), ctypes.byref(nz), ctypes.byref(out), None)
        if rc == L.CRV_E_CAPACITY:
            cap = out.n
            continue
        ctx.check(rc)
        break
    n = out.n
    return {"chan": _host(bufs["label"], n), "time": _host(bufs["time"], n), "charge": _host(bufs["charge"], n),
            "pe": _host(bufs["pe"], n), "photon": _host(bufs["photon"], n)}
