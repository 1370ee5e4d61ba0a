"""The photon-list SiPM stage (crv_sipm_run: dark noise, trap afterpulses) against the oracle, through the C-ABI.

Both sides key every draw by the entry's identity (a photon: label and list position; dark discharge k of a SiPM:
label and k) and share the FP64 arithmetic, so tables agree bit for bit. The oracle itself is pinned to the
reference's sipm.simulate on shared uniforms (tests/test_oracle_reference.py, fixture sipm_noise_vectors.npz).
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import config as CF, sipm as SI  # noqa: E402
from oracle import response as ORS  # noqa: E402

SEED = 4242


def _oracle(chan, t, photon, pmap, cfg, t_start, t_end, thermal):
    chan = np.asarray(chan, np.int64)
    src = ORS.PhiloxSource(SEED, (chan // 32768, chan % 32768, np.arange(len(chan)), np.zeros(len(chan), np.int64)))
    return ORS.sipm_simulate(chan, t, photon, pmap, cfg, src, t_end=t_end, t_start=t_start,
                             thermal_labels=thermal if cfg.thermal_rate > 0 else None)


def _same(Q, R):
    assert len(Q["chan"]) == len(R["chan"])
    for k in ("chan", "photon", "time", "pe", "charge"):
        np.testing.assert_array_equal(Q[k], R[k], err_msg=k)


def _cfg(g):
    return CF.SiPMConfig(**{k: float(g(k)) for k in ("cross_talk", "thermal_rate", "trap0_prob", "trap1_prob",
                                                       "trap0_lifetime", "trap1_lifetime")},
                         pixel_id=str(g("pixel_id")))


@pytest.mark.parametrize("name", ["d", "e", "f"])
def test_noise_cases_match_oracle(noise_golden, name):
    """The inputs of the reference's dark-noise (d), afterpulse (e) and combined (f) fixtures, with labels carrying
    a large event key (the upper word of the Philox key)."""
    g = lambda k: noise_golden[f"noise_{name}_{k}"]
    cfg = _cfg(g)
    off = (3 << 33) * 32768
    chan, th = g("chan_in") + off, g("thermal") + off
    args = (chan, g("time_in"), g("photon_in"), g("pmap"), cfg)
    kw = dict(t_start=float(g("t_start")), t_end=float(g("t_end")))
    Q = SI.simulate(*args, key=SEED, thermal_chans=th, **kw)
    R = _oracle(*args, kw["t_start"], kw["t_end"], th)
    _same(Q, R)
    assert len(Q["chan"]) > 500
    if cfg.thermal_rate > 0:
        assert (Q["photon"] == -1).sum() > 20
    if cfg.trap0_prob > 0:
        Q0 = SI.simulate(*args[:4], dataclasses.replace(cfg, trap0_prob=0.0, trap1_prob=0.0), key=SEED,
                         thermal_chans=th, **kw)
        assert len(Q0["chan"]) < len(Q["chan"])  # afterpulses add charges; the rest is the same simulation
        assert np.isin(Q0["time"], Q["time"]).all()


def test_without_noise_equals_oracle_and_drops_late_photons():
    rng = np.random.default_rng(5)
    chan = np.repeat(np.arange(200) * 11 + 1, rng.poisson(25, 200))
    t = rng.uniform(0, 3000, len(chan))
    photon = rng.integers(0, 10 ** 9, len(chan))
    cfg = CF.SiPMConfig(cross_talk=0.15)
    pmap = rng.random((40, 40))
    Q = SI.simulate(chan, t, photon, pmap, cfg, key=SEED, t_end=2000.0)
    _same(Q, _oracle(chan, t, photon, pmap, cfg, 0.0, 2000.0, None))
    assert Q["time"].max() <= 2000.0
    o = np.lexsort((Q["time"], Q["chan"]))
    np.testing.assert_array_equal(o, np.arange(len(o)))


def test_dark_noise_moments_and_window():
    """Reference crv/tests/test_sipm_jax_thermal.py:107-121 re-expressed: no photons, rate 0.01 per ns over 100 ns,
    no cross-talk and no dead pixels: Poisson(1) fired dark discharges per SiPM, inside the window, PE in (0, 1]."""
    cfg = CF.SiPMConfig(thermal_rate=0.01, cross_talk=0.0, inactive=())
    labels = np.arange(4096)
    Q = SI.simulate([], [], [], np.ones((40, 40)), cfg, key=SEED, t_end=100.0, thermal_chans=labels)
    counts = np.bincount(Q["chan"], minlength=len(labels))
    assert abs(counts.mean() - 1.0) < 5 / np.sqrt(len(labels))
    assert abs(counts.var() - 1.0) < 5 * np.sqrt(3 / len(labels))
    assert np.all(Q["photon"] == -1)
    assert np.all((Q["time"] >= 0) & (Q["time"] <= 100))
    assert np.all((Q["pe"] > 0) & (Q["pe"] <= 1))
    _same(Q, _oracle(np.zeros(0, np.int64), np.zeros(0), np.zeros(0, np.int64), np.ones((40, 40)), cfg, 0.0, 100.0,
                     labels))


def test_zero_rate_or_empty_window_adds_nothing():
    rng = np.random.default_rng(9)
    chan = np.repeat(np.arange(50), 20)
    t = rng.uniform(0, 500, len(chan))
    cfg = CF.SiPMConfig()
    Q0 = SI.simulate(chan, t, np.arange(len(chan)), np.ones((40, 40)), cfg, key=SEED, t_end=1000.0)
    Q1 = SI.simulate(chan, t, np.arange(len(chan)), np.ones((40, 40)), cfg, key=SEED, t_end=1000.0,
                     thermal_chans=np.arange(60))  # rate 0: thermal_chans ignored
    _same(Q0, Q1)
    Q2 = SI.simulate([], [], [], np.ones((40, 40)), dataclasses.replace(cfg, thermal_rate=0.1), key=SEED,
                     t_start=10.0, t_end=10.0, thermal_chans=[0, 1])
    assert all(len(v) == 0 for v in Q2.values())


def test_certain_traps_release_one_afterpulse_per_firing():
    """trap0_prob = p_fire(Vov): every firing is captured. One photon per SiPM at t = 0, no cross-talk: each SiPM
    fires once and its afterpulse comes an Exp(50 ns) delay later and fires with p_fire(v(t)) of the recovering
    pixel. The fired fraction and the mean delay of fired afterpulses follow from that law."""
    cfg0 = CF.SiPMConfig(cross_talk=0.0, inactive=())
    pf = cfg0.p1 * (1 - np.exp(-cfg0.overvoltage / cfg0.p2))
    cfg = dataclasses.replace(cfg0, trap0_prob=pf, trap0_lifetime=50.0)
    n = 20000
    Q = SI.simulate(np.arange(n), np.zeros(n), np.arange(n), np.ones((40, 40)), cfg, key=SEED, t_end=1e4)
    late = Q["time"] > 0
    first = ~late
    nfire = int(first.sum())
    assert abs(nfire - n * pf) < 5 * np.sqrt(n * pf * (1 - pf))
    t = np.linspace(0, 2000, 200001)
    dens = np.exp(-t / 50.0) / 50.0
    v = cfg.overvoltage * (1 - np.exp(-t / cfg.time_constant))
    pfv = cfg.p1 * (1 - np.exp(-v / cfg.p2))
    frac = np.trapezoid(dens * pfv, t)
    mean = np.trapezoid(t * dens * pfv, t) / frac
    k = int(late.sum())
    assert abs(k - nfire * frac) < 5 * np.sqrt(nfire * frac * (1 - frac))
    assert abs(Q["time"][late].mean() - mean) < 5 * 50.0 / np.sqrt(k)
    np.testing.assert_array_equal(np.unique(Q["photon"][late]), np.unique(Q["chan"][late]))  # parent's photon id
