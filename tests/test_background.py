"""Background pulses (crvhip.background, reference reco/background.py): the sampler's laws and the merge order on the
CPU. The reference module is not importable here (it needs awkward), so the sampler is pinned by what its
construction fixes: every sampled pulse is a library pulse shifted by its group's start time, group counts are
Poisson (or gamma-mixed Poisson) with the configured mean, start times are uniform in the window."""
from dataclasses import replace

import numpy as np
import pytest

from crvhip import background as BG, chain as CH, config as CF


def _library(rng, groups=300):
    sizes = rng.integers(1, 6, groups)
    g = np.repeat(np.arange(groups), sizes)
    n = len(g)
    lib = {"group": g, "bar": rng.integers(0, 5000, n), "sipm": rng.integers(0, 4, n), "pe": rng.uniform(1, 40, n),
           "dt": np.concatenate([np.r_[0.0, np.sort(rng.uniform(0, 60, k - 1))] for k in sizes]),
           "seq": rng.integers(0, 2, n)}
    part = dict(events=20000, exposure_ns=20000 * 1e5, groups=16000, cnt_sum=16000, cnt_sumsq=16000 + 1.3 * 16000,
                lib=lib, file="data.root", keep_frac=0.01)
    return BG.BackgroundModel.from_partials([part]), lib


def test_from_partials_rate_overdispersion_and_library():
    M, lib = _library(np.random.default_rng(1))
    assert M.rate_per_ns == pytest.approx(16000 / (20000 * 1e5))
    mu, var = 16000 / 20000, (16000 + 1.3 * 16000) / 20000 - (16000 / 20000) ** 2
    assert M.gamma_k == pytest.approx(mu * mu / (var - mu))
    assert M.count.sum() == len(lib["group"]) and len(M.count) == 300
    np.testing.assert_array_equal(M.start, np.r_[0, np.cumsum(M.count)[:-1]])
    data = dict(events=20000, exposure_ns=2e9, groups=16000, cnt_sum=16000, cnt_sumsq=32000, lib=lib)
    mc = dict(events=1, exposure_ns=1e9, groups=1000, cnt_sum=0, cnt_sumsq=0, lib={k: [] for k in BG.LIB_KEYS})
    assert BG.BackgroundModel.from_partials([data], [mc]).rate_per_ns == pytest.approx(16000 / 2e9 - 1000 / 1e9)
    assert BG.BackgroundModel.from_partials([data], [dict(mc, groups=10 ** 9)]).rate_per_ns == 0.0


def _single_pulse_model(rng, mu_data, k_data, groups=500, window_ns=1300.0):
    """A library of one-pulse groups whose data had mu_data groups per event in a window_ns window with gamma shape
    k_data (group-count variance mu + mu^2 / k)."""
    lib = {"group": np.arange(groups), "bar": rng.integers(0, 5000, groups), "sipm": rng.integers(0, 4, groups),
           "pe": rng.uniform(1, 40, groups), "dt": np.zeros(groups), "seq": rng.integers(0, 2, groups)}
    ev = 100000
    part = dict(events=ev, exposure_ns=ev * window_ns, groups=mu_data * ev, cnt_sum=mu_data * ev,
                cnt_sumsq=(mu_data + mu_data ** 2 / k_data + mu_data ** 2) * ev, lib=lib)
    return BG.BackgroundModel.from_partials([part])


@pytest.mark.parametrize("overdispersed", [False, True])
def test_group_counts_are_poisson_or_gamma_poisson(overdispersed):
    M = _single_pulse_model(np.random.default_rng(2), 3.0, 2.0)
    assert M.gamma_k == pytest.approx(2.0)
    ev = np.arange(200000) * 3 + 11
    B = M.sample(ev, (200.0, 1500.0), np.random.default_rng(5), rate_scale=1.0, overdispersed=overdispersed)
    counts = np.bincount(np.searchsorted(ev, B["ev"]), minlength=len(ev))
    mu = 3.0
    var = mu + mu * mu / 2.0 if overdispersed else mu
    assert abs(counts.mean() - mu) < 5 * np.sqrt(var / len(ev))
    assert abs(counts.var() - var) < 0.05 * var
    assert np.all((B["time"] >= 200.0) & (B["time"] < 1500.0))
    assert abs(B["time"].mean() - 850.0) < 5 * 1300 / np.sqrt(12 * len(B["time"]))


def test_sampled_groups_are_library_groups_at_uniform_starts():
    """Multi-pulse groups: every sampled pulse is a library pulse, the pulses of a draw stay together in library order,
    and time - dt (the group's start) is one value per draw, inside the window."""
    M, lib = _library(np.random.default_rng(3))
    M = replace(M, rate_per_ns=5e-4)
    B = M.sample(np.arange(3000), (200.0, 1500.0), np.random.default_rng(8), overdispersed=False)
    row_of = {(int(b), int(s), float(p)): i for i, (b, s, p) in enumerate(zip(M.bar, M.sipm, M.pe))}
    rows = np.array([row_of[(int(b), int(s), float(p))] for b, s, p in zip(B["bar"], B["sipm"], B["pe"])])
    np.testing.assert_array_equal(B["seq"], M.seq[rows])
    t0 = B["time"] - M.dt[rows]
    grp = np.searchsorted(M.start, rows, side="right") - 1  # library group of each row
    new = np.r_[True, (rows[1:] != rows[:-1] + 1) | (grp[1:] != grp[:-1])]
    draw = np.cumsum(new) - 1
    assert np.all(rows[new] == M.start[grp[new]])          # a draw starts at its group's first pulse
    np.testing.assert_array_equal(np.bincount(draw), M.count[grp[new]])
    at = np.flatnonzero(new)
    assert np.all(np.maximum.reduceat(t0, at) - np.minimum.reduceat(t0, at) < 1e-9)  # one start per draw
    assert np.all((t0 >= 200.0) & (t0 < 1500.0))


def test_no_model_or_zero_scale_samples_nothing():
    M, _ = _library(np.random.default_rng(3))
    assert len(M.sample(np.arange(10), (0, 100), np.random.default_rng(1), rate_scale=0)["ev"]) == 0
    assert BG.sampler(CF.ChainConfig(), np.random.default_rng(1)) is None
    assert BG.sampler(replace(CF.ChainConfig(), background=M, background_scale=0.0), np.random.default_rng(1)) is None


def test_merge_order_and_flags():
    rng = np.random.default_rng(4)
    n = 200
    ev = np.sort(rng.integers(0, 20, n))
    sig = dict(bar=rng.integers(0, 50, n), sipm=rng.integers(0, 4, n), time=rng.uniform(0, 1000, n),
               pe=rng.uniform(1, 30, n), seq=rng.integers(0, 3, n))
    B = {"ev": rng.integers(0, 20, 50), "bar": rng.integers(0, 50, 50), "sipm": rng.integers(0, 4, 50),
         "time": rng.uniform(0, 1000, 50), "pe": rng.uniform(1, 30, 50), "seq": np.zeros(50, np.int64)}
    idx = np.arange(n) - np.searchsorted(ev, ev)
    e, i, b, s, t, p, q, bg = BG.merge_pulses(ev, idx, sig["bar"], sig["sipm"], sig["time"], sig["pe"], sig["seq"], B)
    assert len(e) == n + 50 and bg.sum() == 50
    np.testing.assert_array_equal(np.lexsort((t, s, b, e)), np.arange(len(e)))
    np.testing.assert_array_equal(i, np.arange(len(e)) - np.searchsorted(e, e))
    assert sorted(p[bg]) == sorted(B["pe"])


def test_save_load_round_trip(tmp_path):
    M, _ = _library(np.random.default_rng(6))
    M.save(tmp_path / "bg.npz")
    L = BG.BackgroundModel.load(tmp_path / "bg.npz")
    for k in ("start", "count", "bar", "sipm", "pe", "dt", "seq"):
        np.testing.assert_array_equal(getattr(L, k), getattr(M, k))
    assert L.rate_per_ns == M.rate_per_ns and L.info["library_groups"] == 300


def test_one_call_chains_refuse_a_background():
    M, _ = _library(np.random.default_rng(7))
    cfg = replace(CF.ChainConfig(), background=M)
    with pytest.raises(ValueError, match="background"):
        CH.run_chain({"event": np.zeros(1, int), "counter": np.zeros(1, int), "edep": np.ones(1), "z": np.zeros(1),
                      "t": np.zeros(1)}, None, cfg)
