"""Argument checks of crvhip.sipm.simulate, the SiPM stage with dark noise and trap afterpulses (CPU: every check runs
before the device is touched). Re-expresses the reference's crv/tests/test_sipm_jax_thermal.py:220-266 invalid-input
cases; traps, which the reference's device path refuses, are accepted here (its numpy sipm.simulate runs them), so
their own settings are checked instead. The oracle's keyed dark noise and afterpulses are checked against the laws
they sample (Poisson counts, uniform times and pixels, exponential release delays)."""
from dataclasses import replace

import numpy as np
import pytest

from crvhip import config as CF, sipm as SI
from oracle import philox as PX, response as ORS

PMAP = np.ones((40, 40))


def _sim(cfg=None, **kw):
    args = dict(chan=[], time=[], photon=[])
    args.update({k: kw.pop(k) for k in ("chan", "time", "photon") if k in kw})
    return SI.simulate(args["chan"], args["time"], args["photon"], PMAP, cfg or CF.SiPMConfig(thermal_rate=0.01),
                       key=1, **kw)


@pytest.mark.parametrize("rate", [-1, np.nan, np.inf])
def test_invalid_rates(rate):
    with pytest.raises(ValueError, match="thermal_rate"):
        _sim(CF.SiPMConfig(thermal_rate=rate), thermal_chans=[0])


@pytest.mark.parametrize("kwargs", [dict(t_start=-1), dict(t_end=-1), dict(t_start=2, t_end=1),
                                    dict(t_start=np.nan), dict(t_end=np.inf)])
def test_invalid_windows(kwargs):
    with pytest.raises(ValueError, match="t_start"):
        _sim(thermal_chans=[0], **kwargs)


@pytest.mark.parametrize("labels", [None, [], [0, 0], [-1], [np.nan], [np.inf], [1.5], [[1]], [True],
                                    np.array([2 ** 63], dtype=np.uint64)])
def test_invalid_thermal_labels(labels):
    with pytest.raises(ValueError, match="thermal_chans"):
        _sim(thermal_chans=labels)


@pytest.mark.parametrize("changes", [dict(n_pixels_x=0), dict(n_pixels_y=42), dict(n_pixels_x=2.5),
                                     dict(n_pixels_x=True), dict(inactive=((41, 0),)),
                                     dict(inactive=((-1, 0),)), dict(inactive=((1.5, 0),))])
def test_invalid_pixels(changes):
    with pytest.raises(ValueError, match="pixel"):
        _sim(replace(CF.SiPMConfig(thermal_rate=0.1), **changes), thermal_chans=[0])


@pytest.mark.parametrize("changes", [dict(trap0_prob=-1), dict(trap1_prob=np.nan), dict(trap0_prob=0.1, trap0_lifetime=0),
                                     dict(trap1_prob=0.1, trap1_lifetime=np.inf)])
def test_invalid_traps(changes):
    with pytest.raises(ValueError, match="trap"):
        _sim(replace(CF.SiPMConfig(), **changes))


@pytest.mark.parametrize("p1", [0, -1, np.nan, np.inf, 10])
def test_invalid_thermal_firing_probability(p1):
    with pytest.raises(ValueError, match="firing probability"):
        _sim(CF.SiPMConfig(thermal_rate=0.1, p1=p1), thermal_chans=[0])


@pytest.mark.parametrize("chan,time,photon", [([-1], [0], [0]), ([0.5], [0], [0]), ([0], [np.nan], [0]),
                                             ([0], [], [0]), ([[0]], [0], [0])])
def test_invalid_photons(chan, time, photon):
    with pytest.raises(ValueError):
        _sim(CF.SiPMConfig(), chan=chan, time=time, photon=photon)


def test_unknown_keyword():
    with pytest.raises(TypeError, match="unexpected keyword"):
        _sim(thermal_channels=[0])


def test_keyed_dark_noise_follows_its_laws():
    """PhiloxSource.thermal: Poisson(lam) counts per label (mean and variance), uniform times over [t_start, t_end)
    and uniform pixels over the grid (each coordinate's bins within 5 sigma), distinct identities."""
    labels = np.arange(20000) * 3 + (7 << 33) * 32768
    src = ORS.PhiloxSource(11, (np.zeros(0, np.int64),) * 4)
    lam = 2.5
    tl, tt = src.thermal(labels, lam, 100.0, 1500.0)
    counts = np.unique(tl, return_counts=True)[1]
    counts = np.r_[counts, np.zeros(len(labels) - len(counts), np.int64)]
    assert abs(counts.mean() - lam) < 5 * np.sqrt(lam / len(labels))
    assert abs(counts.var() - lam) < 5 * np.sqrt((lam + 2 * lam ** 2) / len(labels))
    assert tt.min() >= 100.0 and tt.max() < 1500.0
    assert abs(tt.mean() - 800.0) < 5 * 1400 / np.sqrt(12 * len(tt))
    px, py = src.dark_pixels(len(tl), 40, 40)
    for c in (px, py):
        assert c.min() == 0 and c.max() == 39
        b = np.bincount(c, minlength=40)
        assert np.max(np.abs(b - len(c) / 40)) < 5 * np.sqrt(len(c) / 40)
    ids = np.stack([src.ev, src.ch, src.j, src.r], 1)
    assert len(np.unique(ids, axis=0)) == len(tl) and np.all(src.j == PX.DARK_J)


def test_keyed_afterpulses_follow_their_law():
    """Every firing captured (trap0_prob = p_fire(Vov)), one photon per SiPM at t = 0, no cross-talk: the afterpulse
    of a SiPM fires with the integral of Exp(50 ns) x p_fire(v(t)) of its recovering pixel."""
    cfg0 = CF.SiPMConfig(cross_talk=0.0, inactive=())
    pf = cfg0.p1 * (1 - np.exp(-cfg0.overvoltage / cfg0.p2))
    cfg = replace(cfg0, trap0_prob=pf, trap0_lifetime=50.0)
    n = 6000
    chan = np.arange(n)
    src = ORS.PhiloxSource(3, (chan // 32768, chan % 32768, chan, np.zeros(n, np.int64)))
    Q = ORS.sipm_simulate(chan, np.zeros(n), chan, PMAP, cfg, src, t_end=1e4)
    late = Q["time"] > 0
    nfire = int((~late).sum())
    t = np.linspace(0, 2000, 200001)
    dens = np.exp(-t / 50.0) / 50.0
    pfv = cfg.p1 * (1 - np.exp(-cfg.overvoltage * (1 - np.exp(-t / cfg.time_constant)) / cfg.p2))
    frac = np.trapezoid(dens * pfv, t)
    k = int(late.sum())
    assert abs(k - nfire * frac) < 5 * np.sqrt(nfire * frac * (1 - frac))
    assert set(Q["photon"][late]) <= set(Q["photon"][~late])  # an afterpulse carries its parent's photon id
