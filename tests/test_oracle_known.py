"""The oracle against known answers: Random123 vectors, closed forms, and the reference tests' expectations.

Each test cites the reference test it re-expresses (paths relative to the reference repository) or the mathematics
it checks, chosen so that a dropped term, a wrong index or a swapped operand fails.
"""
from dataclasses import replace

import numpy as np
import pytest
from scipy import stats

from oracle import coinc as OCo, fmath as F, philox as PX, pulsefit as OPF, rand as R, response as ORS
from oracle import waveform as OWF
from crvhip import config as CF, detector as DT, template as TP


# --- Philox4x32-10 (Random123 kat_vectors) ---------------------------------------------------------------------------

@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    out = PX.philox4x32(*[np.array([c], np.uint32) for c in ctr], np.array([key[0]], np.uint32),
                        np.array([key[1]], np.uint32))
    assert tuple(int(o[0]) for o in out) == expect


def test_uniforms_open_interval_and_exact():
    u = PX.uniform(np.array([0, 1, 0xFFFFFFFF], np.uint32))
    assert u[0] == 2.0 ** -33 and u[2] == 1.0 - 2.0 ** -33 and np.all((u > 0) & (u < 1))


# --- portable transcendentals -----------------------------------------------------------------------------------------

def test_fmath_close_to_libm():
    rng = np.random.default_rng(0)
    x = rng.uniform(-700, 700, 200_000)
    assert np.max(np.abs(F.exp(x) - np.exp(x)) / np.spacing(np.exp(x))) <= 2
    x = np.exp(rng.uniform(-700, 700, 200_000))
    assert np.max(np.abs(F.log(x) - np.log(x)) / np.spacing(np.abs(np.log(x)))) <= 3
    x = rng.uniform(-500, 500, 200_000)
    assert np.max(np.abs(F.sin(x) - np.sin(x))) <= 4e-16 and np.max(np.abs(F.cos(x) - np.cos(x))) <= 4e-16
    x = rng.uniform(-1, 1, 200_000)
    assert np.max(np.abs(F.asin(x) - np.arcsin(x)) / np.spacing(np.abs(np.arcsin(x)))) <= 4


def test_fmath_special_values():
    e = F.exp(np.array([0.0, -np.inf, np.inf, -800.0, 800.0, 1.0]))
    assert e[0] == 1.0 and e[1] == 0.0 and e[2] == np.inf and e[3] == 0.0 and e[4] == np.inf
    assert abs(e[5] - np.e) <= np.spacing(np.e)
    l = F.log(np.array([0.0, 1.0, np.inf, 2.0]))
    assert l[0] == -np.inf and l[1] == 0.0 and l[2] == np.inf and abs(l[3] - np.log(2)) <= np.spacing(1.0)
    assert np.isnan(F.log(np.array([-1.0])))[0]
    assert F.asin(np.array([1.0]))[0] == F.PI_2 and F.asin(np.array([0.0]))[0] == 0.0
    assert F.PIO2_1 + F.PIO2_2 + F.PIO2_3 == np.pi / 2


# --- samplers ---------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("mu", [0.3, 4.0, 9.99, 10.0, 37.5, 400.0])
def test_poisson_moments_and_pmf(mu):
    n = 200_000
    ev = np.arange(n)
    k, capped = R.poisson(np.full(n, mu), lambda call: PX.draw(3, PX.KIND_COUNT, ev, 5, 1, call))
    assert not capped.any()
    se = np.sqrt(mu / n)
    assert abs(k.mean() - mu) < 5 * se
    assert abs(k.var() / mu - 1) < 5 * np.sqrt(2.0 / n) + 0.01
    # chi2 goodness of fit against the Poisson pmf on the central bins
    lo, hi = stats.poisson.ppf([0.001, 0.999], mu)
    bins = np.arange(lo, hi + 1)
    obs = np.array([(k == b).sum() for b in bins])
    exp_ = stats.poisson.pmf(bins, mu) * n
    good = exp_ > 20
    chi2 = ((obs[good] - exp_[good]) ** 2 / exp_[good]).sum()
    assert stats.chi2.sf(chi2, good.sum() - 1) > 1e-4


def test_poisson_zero_mean():
    k, _ = R.poisson(np.zeros(10), lambda call: PX.draw(1, PX.KIND_COUNT, np.arange(10), 0, 0, call))
    assert np.all(k == 0)


def test_loggam_matches_lgamma():
    from math import lgamma
    x = np.arange(1, 3000, dtype=float)
    got = R.loggam_int(x)
    want = np.array([lgamma(v) for v in x])
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1)) < 5e-15


def test_ordered_delays_are_sorted_exponential_order_statistics():
    tau, N, G = 7.4, 50, 4000
    S = np.array([R.ordered_delays(N, tau, lambda call, g=g: PX.draw(2, PX.KIND_DELAY, g, 1, 0, call))
                  for g in range(G)])
    assert np.all(np.diff(S, axis=1) >= 0)
    # the k-th order statistic of N Exp(tau) has mean tau * sum_{i<k} 1/(N - i)
    ks = np.array([0, 9, 24, 49])
    mean_th = tau * np.cumsum(1.0 / (N - np.arange(N)))[ks]
    assert np.all(np.abs(S[:, ks].mean(0) / mean_th - 1) < 0.05)
    # pooled delays are Exp(tau)
    assert stats.kstest(S.ravel()[::7], stats.expon(scale=tau).cdf).pvalue > 1e-3


def test_box_muller_normal():
    u = PX.draw(4, PX.KIND_NOISE, np.arange(100_000), 2, 3, 0)
    z = R.normal(u[0], u[1])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-3


# --- photons: closed-form yield -------------------------------------------------------------------------------------

def test_photon_yield_follows_two_exponential_attenuation():
    opt = CF.OpticsConfig()
    n = 4000
    z = np.repeat([-2000.0, 0.0, 2000.0], n)
    steps = dict(evkey=np.arange(3 * n), counter=np.zeros(3 * n, np.int64), edep=np.full(3 * n, 2.0), z=z,
                 t=np.zeros(3 * n))
    P = ORS.photons(steps, np.zeros(4 * DT.N_COUNTERS, np.int64), np.ones(4 * DT.N_COUNTERS), opt, seed=1)
    A = lambda d: 0.2 * np.exp(-d / 300.0) + 0.8 * np.exp(-d / 4500.0)
    for zi in (-2000.0, 0.0, 2000.0):
        for s in range(4):
            d = zi + 2550.0 if s % 2 == 0 else 2550.0 - zi
            mu = 39.0 * 2.0 * A(d) / A(1000.0)
            sel = (P["chan"] == s) & (steps["z"][P["row"]] == zi)
            assert abs(sel.sum() / n - mu) < 5 * np.sqrt(mu / n), (zi, s)
    # arrival = t + d / v + delay >= t + d / v, and the delays of one (step, SiPM) come sorted
    d = np.where(P["chan"] % 2 == 0, steps["z"][P["row"]] + 2550.0, 2550.0 - steps["z"][P["row"]])
    assert np.all(P["time"] >= d / 140.0 - 1e-12)
    grp = P["row"] * 4 + P["chan"]
    same = grp[1:] == grp[:-1]
    assert np.all(np.diff(P["time"])[same] >= 0)


def test_status_mask_drops_bits_0_and_2():
    opt = CF.OpticsConfig()
    status = np.zeros(4 * DT.N_COUNTERS, np.int64)
    status[[0, 2]] = [1, 4]
    status[1] = 2                        # bit 1 only ignores pulses later, not the SiPM
    steps = dict(evkey=np.arange(50), counter=np.zeros(50, np.int64), edep=np.full(50, 3.0), z=np.zeros(50),
                 t=np.zeros(50))
    P = ORS.photons(steps, status, np.ones(len(status)), opt, seed=2)
    assert set(np.unique(P["chan"])) == {1, 3}


# --- SiPM: reference crv/tests/test_photons.py:82-104 -----------------------------------------------------------------

def test_sipm_single_photons_full_pe_and_crosstalk_rate():
    cfg = CF.SiPMConfig(cross_talk=0.04)
    n = 200_000
    src = ORS.PhiloxSource(3, (np.arange(n), np.zeros(n), np.zeros(n), np.zeros(n)))
    out = ORS.sipm_simulate(np.arange(n), np.zeros(n), np.arange(n), np.ones((40, 40)), cfg, src)
    assert np.allclose(out["pe"], 1.0)
    p = cfg.p1 * (1 - np.exp(-cfg.overvoltage / cfg.p2))
    live = 1 - 0.1 * 0.1
    ct = 4 * (1 - (1 - cfg.cross_talk) ** 0.25)
    assert abs(len(out["pe"]) / n - p * live * (1 + ct)) < 0.004


def test_sipm_recovery_reduces_second_firing():
    cfg = CF.SiPMConfig(cross_talk=0.0)
    pmap = np.zeros((40, 40))
    pmap[5, 5] = 1.0
    n = 50_000
    lab = np.repeat(np.arange(n), 2)
    src = ORS.PhiloxSource(4, (lab, np.zeros(2 * n), np.zeros(2 * n), np.tile([0, 1], n)))
    out = ORS.sipm_simulate(lab, np.tile([0.0, 1.0], n), np.arange(2 * n), pmap, cfg, src)
    late = out["pe"][out["time"] == 1.0]
    v1 = 1 - np.exp(-1 / cfg.time_constant)
    assert np.isclose(late.min(), v1) and np.isclose(late.max(), 1.0)


def test_sipm_brute_force_two_photons_same_pixel():
    """Two photons on one pixel 2 ns apart, no cross talk: P(second fires) = P1 * p(v(2ns)) + (1 - P1) * p(Vov)."""
    cfg = CF.SiPMConfig(cross_talk=0.0)
    pmap = np.zeros((40, 40))
    pmap[5, 5] = 1.0
    n = 100_000
    lab = np.repeat(np.arange(n), 2)
    src = ORS.PhiloxSource(8, (lab, np.zeros(2 * n), np.zeros(2 * n), np.tile([0, 1], n)))
    out = ORS.sipm_simulate(lab, np.tile([0.0, 2.0], n), np.arange(2 * n), pmap, cfg,
                            src)
    pf = lambda v: cfg.p1 * (1 - np.exp(-v / cfg.p2))
    # lrint: pixel (5|6, 5|6) with equal odds: same pixel w.p. 1/4
    p_full = pf(cfg.overvoltage)
    p_rec = pf(cfg.overvoltage * (1 - np.exp(-2.0 / cfg.time_constant)))
    p_second = 0.25 * (p_full * p_rec + (1 - p_full) * p_full) + 0.75 * p_full
    got = (out["time"] == 2.0).sum() / n
    assert abs(got - p_second) < 5 * np.sqrt(p_second / n)


def test_sipm_pixel_ids_lrint_and_truncate():
    """Reference crv/tests/test_sipm_pixel_id.py:21-31: lrint is the bin or the bin + 1 with odds 1/2 per axis."""
    rng = np.random.default_rng(1)
    pmap = np.zeros((40, 40))
    pmap[6:34, 6:34] = rng.random((28, 28))
    cum = ORS.cdf_of_map(pmap)
    u = [rng.random(200_000) for _ in range(3)]
    tx, ty = ORS.pixel_ids(cum, 40, *u, "truncate")
    lx, ly = ORS.pixel_ids(cum, 40, *u, "lrint")
    assert tx.min() == 6 and tx.max() == 33 and np.all(pmap[tx, ty] > 0)
    dx = lx - tx
    assert set(np.unique(dx)) == {0, 1} and abs((dx == 0).mean() - 0.5) < 0.01 and lx.max() == 34


# --- waveforms: reference crv/tests/test_reco_pulsefit.py:66-93 -------------------------------------------------------

def _tmpl():
    return TP.load(CF.WaveformConfig())


def test_waveform_linear_and_zero_suppressed():
    cfg = CF.WaveformConfig()
    t = 300.0 + np.linspace(0.0, 6.0, 30)
    q = np.full(30, cfg.single_pe_reference_charge)
    chan = np.zeros(30, np.int64)
    a = OWF.synthesize(chan, t, q, np.array([3.0]), _tmpl(), cfg)
    b = OWF.synthesize(chan, t, 2 * q, np.array([3.0]), _tmpl(), cfg)
    assert len(a["chan"]) >= 1
    assert np.allclose(b["volt"][0], 2 * a["volt"][0])
    assert np.allclose((a["start"] - 3.0) / 12.5, np.rint((a["start"] - 3.0) / 12.5))
    tiny = OWF.synthesize(chan[:1], t[:1], np.array([0.5 * cfg.single_pe_reference_charge]), np.array([3.0]),
                          _tmpl(), cfg)
    assert len(tiny["chan"]) == 0


def test_waveform_then_reco_roundtrip():
    cfg = CF.WaveformConfig()
    rng = np.random.default_rng(3)
    nq = 40
    t = np.sort(300.0 + rng.exponential(8.0, nq))
    q = np.full(nq, cfg.single_pe_reference_charge)
    dg = OWF.synthesize(np.zeros(nq, np.int64), t, q, np.array([0.0]), _tmpl(), cfg)
    lens = np.full(len(dg["chan"]), cfg.samples_zs)
    gid, mat, n, start, first = OWF.concat_waveforms(np.zeros(len(lens), np.int64), dg["chan"],
                                                     np.zeros(len(lens), np.int64), dg["starttdc"],
                                                     dg["adc"].ravel(), lens)
    res = OPF.reconstruct(mat, n, start, np.full(len(n), 100.0), np.ones(len(n)), np.ones(len(n)), OPF.RecoConfig())
    assert len(res["wf"]) >= 1 and 290.0 < res["time"][0] < 360.0


def test_waveform_window_cap_counts_overlong_clusters():
    cfg = CF.WaveformConfig()
    t = np.arange(0.0, 2000.0, 100.0)          # gaps of 100 ns < 150 ns: one cluster 1,900 ns long
    q = np.full(len(t), 20 * cfg.single_pe_reference_charge)
    D = OWF.synthesize(np.zeros(len(t), np.int64), t, q, np.array([0.0]), _tmpl(), cfg, window_samples=128)
    assert D["overlong"] == 1 and D["cluster_length"][0] == 116 and D["record_adc"].shape == (1, 128)
    assert np.all(D["digi_sample"] + 12 <= 128)


def test_adc_saturates_and_matches_int16_cast_in_range():
    cfg = CF.WaveformConfig()
    v = np.array([-1.0, -0.05, 0.0, 0.01, 1.5, 1.78, 50.0])
    adc = OWF.adc_of(v, cfg)
    ref = np.clip((v * 2300 + 100 + 0.5).astype(np.int16), 0, 4095)
    inrange = np.abs(v * 2300) < 30000
    np.testing.assert_array_equal(adc[inrange], ref[inrange])
    assert adc[-1] == 4095


# --- chain glue: reference crv/tests/test_chain.py:18-32 ---------------------------------------------------------------

def _cond16():
    z = np.zeros(16)
    return DT.Conditions(z + 100, z + 400, z + 11, z, np.zeros(16, np.int64), np.ones(16, np.int64),
                         np.repeat(np.arange(4), 4), z)


def test_tdc0_one_draw_per_event_and_feb():
    lab = np.array([1 * 32768 + 0, 1 * 32768 + 3, 1 * 32768 + 4, 2 * 32768 + 0])
    t = OWF.tdc0_per_feb(lab, _cond16(), 1)
    assert t[0] == t[1] and t[0] != t[2] and t[0] != t[3]
    assert np.all((t >= 0) & (t < 12.5))


def test_reco_waveforms_concatenates_consecutive_digis():
    lab = np.array([5 * 32768 + 2] * 3)
    D = {"chan": lab, "starttdc": np.array([10, 22, 40]), "adc": np.arange(36).reshape(3, 12) + 100}
    W = OWF.reco_waveforms(D, _cond16())
    assert W["n"].tolist() == [24, 12] and W["start"].tolist() == [10, 40] and W["ev"].tolist() == [5, 5]
    assert W["bar"].tolist() == [0, 0] and W["sipm"].tolist() == [2, 2] and np.all(W["ped"] == 100)


def test_concat_waveforms_joins_continuing_digis():
    """Reference crv/tests/test_reco_pulsefit.py:53-63."""
    gid, mat, n, start, first = OWF.concat_waveforms(np.array([1, 1, 1, 2]), np.array([5, 5, 6, 5]), np.zeros(4, int),
                                                     np.array([10, 22, 10, 10]), np.arange(48), np.full(4, 12))
    assert list(gid) == [0, 0, 1, 2] and list(n) == [24, 12, 12] and list(start) == [10, 10, 10]
    assert np.array_equal(mat[0, :24], np.arange(24))


# --- pulse reco: reference crv/tests/test_reco_pulsefit.py:23-50, test_chain_reco_contract.py ----------------------------

def _wave(A, mu, b, n=24, start_tdc=10, ped=100.0):
    t = (start_tdc + np.arange(n)) * 12.5
    return np.round(OPF.gumbel(t, A, mu, b) + ped).astype(np.int64)


def test_single_pulse_fit_recovers_parameters():
    W = 200
    rng = np.random.default_rng(1)
    A = rng.uniform(300, 2000, W)
    mu = 16 * 12.5 + rng.uniform(-6, 6, W)
    b = rng.uniform(14, 24, W)
    adc = np.stack([_wave(A[i], mu[i], b[i]) for i in range(W)])
    res = OPF.reconstruct(adc, np.full(W, 24), np.full(W, 10), np.full(W, 100.0), np.ones(W), np.ones(W))
    first = res["seq"] == 0
    assert first.sum() == W
    ok = ~res["failed"][first]
    assert ok.mean() > 0.95
    assert np.median(np.abs(res["fit_mu"][first][ok] - mu[ok])) < 0.3
    assert np.median(np.abs(res["fit_b"][first][ok] / b[ok] - 1)) < 0.03
    assert np.median(np.abs(res["pe"][first][ok] / (A[ok] * b[ok]) - 1)) < 0.03


def test_below_threshold_and_undershoot():
    res = OPF.reconstruct(_wave(30.0, 200.0, 19.0)[None], np.array([24]), np.array([10]), np.array([100.0]),
                          np.ones(1), np.ones(1))
    assert len(res["wf"]) == 0
    adc = _wave(800.0, 200.0, 19.0, ped=90.0)[None]
    res = OPF.reconstruct(adc, np.array([24]), np.array([10]), np.array([100.0]), np.ones(1), np.ones(1))
    assert not res["pedfromdb"][0] and res["ped"][0] == adc[0, 0]


def test_full_length_reco_finds_every_pulse_of_a_train():
    """Reference crv/tests/test_chain_reco_contract.py:120-127: 10 pulses over 252 samples, not capped at 8."""
    t = np.arange(252) * 12.5
    sig = OPF.gumbel(t[None, :], 1000, np.arange(4, 244, 24)[:, None] * 12.5, 12).sum(0)
    adc = np.rint(sig + 100).astype(np.int64)[None]
    res = OPF.reconstruct_full(adc, np.array([252]), np.array([0]), np.array([100.]), np.array([400.]),
                               np.array([11.]))
    assert len(res["pe"]) == 10


def test_solve3_and_jacobi_against_closed_forms():
    rng = np.random.default_rng(0)
    A = rng.normal(size=(500, 3, 3))
    A[:, 0, 0] = 0.0                               # forces a pivot swap
    b = rng.normal(size=(500, 3))
    x, sing = OPF.solve3(A, b)
    assert not sing.any()
    np.testing.assert_allclose(np.einsum("kij,kj->ki", A, x), b, atol=1e-9)
    S = np.einsum("kij,klj->kil", A, A)
    lo, hi = OPF.eig3_minmax(S)
    ev = np.linalg.eigvalsh(S)
    np.testing.assert_allclose(lo, ev[:, 0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(hi, ev[:, -1], rtol=1e-12)
    x, sing = OPF.solve3(np.zeros((1, 3, 3)), np.ones((1, 3)))
    assert sing.all()


# --- coincidences: reference crv/tests/test_coinc_attachment_cpu.py ---------------------------------------------------

def _geom(both=True):
    layer = np.tile(np.arange(4), 2)
    x = np.repeat([0., 400.], 4)
    y = layer * 10.
    return DT.SectorGeom(sector=np.zeros(8, dtype=int), layer=layer, counter_in_type=np.repeat([0, 1], 4), x=x, y=y,
                         pos=np.column_stack([x, y, np.zeros(8)]), length_dir=np.array([2]),
                         sector_type=np.array([1]), sector_names=["EX"], pos_side0=np.array([-700.]),
                         pos_side1=np.array([700.]), sipms_side0=np.array([True]), sipms_side1=np.array([both]))


def _config(**kw):
    return OCo.FinderConfig(cuts={"EX": OCo.SectorCuts(8, 10, 10., -6, 6, 3, 3, 0, 250, 0)}, **kw)


def _pulse(idx, bar, sipm, time=100., pe=10., seq=0):
    return dict(idx=idx, bar=bar, sipm=sipm, time=time, pe=pe, seq=seq)


def _seeds(time=100., start=10, sipm=0):
    return [_pulse(start + bar, bar, sipm, time) for bar in range(3)]


def _run(rows, cfg=None, G=None, status=None):
    p = {key: np.array([row[key] for row in rows]) for key in rows[0]}
    return OCo.find_event(p, _geom() if G is None else G, np.zeros(32, int) if status is None else status,
                          _config() if cfg is None else cfg)


def _members(rows):
    return [row["pulses"] for row in rows]


def test_coinc_attachment_default_and_disabled():
    p = _seeds() + [_pulse(13, 0, 1, 105.)]
    got = _run(p)
    assert _members(got) == [[10, 11, 12, 13]]
    assert got[0]["pe"] == 40. and got[0]["sidehits"] == [3, 1] and got[0]["sidepe"] == [30., 10.]
    assert got[0]["end"] == 105.
    assert _members(_run(p, _config(add_opposite_side_pulses=False))) == [[10, 11, 12]]
    assert _members(_run(p, G=_geom(both=False))) == [[10, 11, 12]]


@pytest.mark.parametrize("time,attached", [
    (np.nextafter(120., -np.inf), True), (120., False), (np.nextafter(120., np.inf), False),
    (np.nextafter(80., np.inf), True), (80., False), (np.nextafter(80., -np.inf), False)])
def test_coinc_strict_attachment_window(time, attached):
    got = _run(_seeds() + [_pulse(13, 0, 1, time)])
    assert _members(got) == [[10, 11, 12] + ([13] if attached else [])]


def test_coinc_filtered_candidates_and_inclusive_filter_boundary():
    assert _members(_run(_seeds() + [_pulse(13, 0, 1, 105., pe=7.)])) == [[10, 11, 12]]
    p = _seeds() + [_pulse(13, 0, 3, 100., pe=4.), _pulse(14, 0, 1, 110., pe=4.)]
    assert _members(_run(p)) == [[10, 11, 12, 14, 13]]
    p[-1]["time"] = np.nextafter(110., np.inf)
    assert _members(_run(p)) == [[10, 11, 12]]


def test_coinc_compensated_pe():
    status = np.zeros(32, int)
    status[3] = 1
    got = _run(_seeds() + [_pulse(13, 0, 1, 105., pe=4.)], status=status)
    assert _members(got) == [[10, 11, 12, 13]] and got[0]["sidepe"] == [30., 8.]


def test_coinc_candidate_tie_and_sorted_bar_sipm_append():
    p = _seeds() + [_pulse(30, 1, 3, 103.), _pulse(40, 0, 3, 104.), _pulse(90, 0, 1, 102.), _pulse(2, 0, 1, 102.),
                    _pulse(50, 1, 1, 101.)]
    assert _members(_run(p)) == [[10, 11, 12, 90, 40, 50, 30]]


def test_coinc_nearest_cluster_tie_and_loser_not_reassigned():
    p = _seeds() + _seeds(time=130., start=20) + [_pulse(30, 0, 1, 115.), _pulse(31, 0, 1, 114.)]
    assert _members(_run(p)) == [[10, 11, 12, 31], [20, 21, 22]]
    assert _members(_run(p[:-1])) == [[10, 11, 12, 30], [20, 21, 22]]
    p[-2]["time"] = np.nextafter(115., np.inf)
    assert _members(_run(p[:-1])) == [[10, 11, 12], [20, 21, 22, 30]]


def test_coinc_secondary_members_occupy_side():
    p = _seeds() + _seeds(start=20, sipm=1) + [_pulse(30, 0, 3, 115.)]
    p[3]["seq"] = 1
    assert _members(_run(p, _config(ignore_secondary_pulses=False))) == [[10, 11, 12, 20, 21, 22]]


def test_coinc_attachment_rescues_min_cluster_pes():
    cfg = OCo.FinderConfig(cuts={"EX": OCo.SectorCuts(8, 10, 10., -6, 6, 3, 3, 35., 250, 0)})
    p = _seeds() + [_pulse(13, 0, 1, 105.)]
    assert _members(_run(p, cfg)) == [[10, 11, 12, 13]]
    assert _run(p, replace(cfg, add_opposite_side_pulses=False)) == []


# --- checkCoincidence: the numpy path equals the combination loop ----------------------------------------------------

def _random_cluster(rng, n_per_layer, mixed):
    cuts = [OCo.SectorCuts(pe_threshold=8.0, max_dt_adjacent=10.0, max_dt=10.0, min_slope=-6.0, max_slope=6.0,
                           max_slope_diff=2.0, coincidence_layers=3, min_cluster_pes=10.0, initial_max_distance=100.0,
                           final_max_distance=60.0),
            OCo.SectorCuts(pe_threshold=8.0, max_dt_adjacent=10.0, max_dt=6.0, min_slope=-2.0, max_slope=3.0,
                           max_slope_diff=1.0, coincidence_layers=2, min_cluster_pes=10.0, initial_max_distance=100.0,
                           final_max_distance=40.0)]
    hits, idx = [], 0
    for layer, k in enumerate(n_per_layer):
        for _ in range(k):
            col = int(rng.integers(0, 6))
            c = cuts[int(rng.integers(0, 2))] if mixed else cuts[0]
            t = float(rng.choice([0.0, 4.0, 9.0, 12.5]) + rng.integers(0, 3) * 0.5)   # repeated times, window edges
            hits.append(dict(idx=idx, layer=layer, x=50.0 * col, y=20.0 * layer, t=t, cuts=c, maxd=100.0))
            idx += 1
    rng.shuffle(hits)
    return hits


@pytest.mark.parametrize("mixed", [False, True])
def test_check_coincidence_numpy_path_equals_combination_loop(mixed, monkeypatch):
    """Hits on few columns (x ties, so the stable sort by x matters), times on a coarse grid (window edges hit
    exactly), and per-hit cuts from two sectors (skip rule, widest window and slope range of a combination)."""
    rng = np.random.default_rng(11 + mixed)
    cfg = OCo.FinderConfig()
    for trial in range(40):
        cl = _random_cluster(rng, [int(v) for v in rng.integers(0, 5, 4)], mixed)
        monkeypatch.setattr(OCo, "NUMPY_COMBOS", 1 << 40)
        loop = OCo.check_coincidence([dict(h) for h in cl], cfg)
        monkeypatch.setattr(OCo, "NUMPY_COMBOS", -1)
        vec = OCo.check_coincidence([dict(h) for h in cl], cfg)
        assert [h["idx"] for h in vec] == [h["idx"] for h in loop], trial
        assert [h["maxd"] for h in vec] == [h["maxd"] for h in loop], trial
