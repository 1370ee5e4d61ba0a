"""crv_pulses_run_noalloc (scratch in one fixed device array, no allocation) against crv_pulses_run through the C ABI:
the same pulse table bit for bit on two calls in a row, the second smaller, and the size limits refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from crvhip import _lib as L, chain as CH, config as CF, detector as DT, synth as SY, template as TP  # noqa: E402
from oracle import chain as OC  # noqa: E402


def _device_waveforms(W, cond):
    n = np.asarray(W["n"], np.int64)
    adc = np.asarray(W["adc"])
    mask = np.arange(adc.shape[1])[None, :] < n[:, None]
    chan = np.asarray(W["bar"]) * 4 + np.asarray(W["sipm"])
    args = [CH._dev(np.asarray(W["ev"], np.int64), np.int64), CH._dev(chan.astype(np.int32), np.int32),
            CH._dev(np.asarray(W["start"], np.int64), np.int64), CH._dev(n.astype(np.int32), np.int32),
            CH._dev(np.r_[0, np.cumsum(n)], np.int64), CH._dev(adc[mask].astype(np.int16), np.int16)]
    cal = [CH._dev(np.asarray(W[k], np.float64), np.float64) for k in ("ped", "area", "height", "toff")]
    return args, cal, len(n), int(n.sum() // 2 + len(n))


def _run(ctx, fn, args, cal, nw, cap):
    b, s = CH._pulse_bufs(cap)
    rc = fn(ctx.h, *[L.ptr(a) for a in args], nw, *[L.ptr(a) for a in cal], ctypes.byref(s), None)
    ctx.check(rc)
    return {k: CH._host(v, s.n) for k, v in b.items()}, s.n


@pytest.fixture(scope="module")
def waves():
    cfg = CF.ChainConfig()
    cond = DT.nominal_conditions()
    tmpl = TP.load(cfg.waveform)
    Qo = OC.charges(SY.uneven_events(120, seed=7), cond, cfg, 2)
    W = OC.reco_waveforms(OC.digis(Qo, cond, cfg, 2, tmpl), cond)
    return cfg, cond, W


def test_noalloc_equals_the_allocating_entry_point_on_two_calls(waves):
    cfg, cond, W = waves
    ctx = CH.context(cfg, cond, DT.sector_geom())
    ctx.set_params(cfg, 0)
    nw_all = len(W["n"])
    small = {k: (v[: nw_all // 5] if isinstance(v, np.ndarray) and len(v) == nw_all else v) for k, v in W.items()}
    for Wc in (W, small):
        args, cal, nw, cap = _device_waveforms(Wc, cond)
        ref, nref = _run(ctx, ctx.lib.crv_pulses_run, args, cal, nw, cap)
        got, ngot = _run(ctx, ctx.lib.crv_pulses_run_noalloc, args, cal, nw, cap)
        assert ngot == nref > 1000
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_noalloc_refuses_calls_beyond_its_scratch(waves):
    cfg, cond, W = waves
    ctx = CH.context(cfg, cond, DT.sector_geom())
    ctx.set_params(cfg, 0)
    args, cal, nw, cap = _device_waveforms(W, cond)
    b, s = CH._pulse_bufs(cap)
    rc = ctx.lib.crv_pulses_run_noalloc(ctx.h, *[L.ptr(a) for a in args], 1 << 17 | 1, *[L.ptr(a) for a in cal],
                                        ctypes.byref(s), None)
    assert rc == L.CRV_E_ARG
    with pytest.raises(L.CrvError, match="exceed the scratch limit"):
        ctx.check(rc)
    # 50,000 waveforms of 12 samples: 600,000 samples, more than the 524,288 the scratch holds
    reps = 50_000 // len(W["n"]) + 1
    big = {k: (np.concatenate([v] * reps)[:50_000] if isinstance(v, np.ndarray) and len(v) == len(W["n"]) else v)
           for k, v in W.items()}
    big["n"] = np.full(50_000, 12)
    big["adc"] = big["adc"][:, :12] if big["adc"].shape[1] >= 12 else np.pad(big["adc"], ((0, 0), (0, 12)))[:, :12]
    args, cal, nw, cap = _device_waveforms(big, cond)
    b, s = CH._pulse_bufs(cap)
    rc = ctx.lib.crv_pulses_run_noalloc(ctx.h, *[L.ptr(a) for a in args], nw, *[L.ptr(a) for a in cal],
                                        ctypes.byref(s), None)
    assert rc == L.CRV_E_ARG
    # the allocating entry point takes the same call
    ctx.check(ctx.lib.crv_pulses_run(ctx.h, *[L.ptr(a) for a in args], nw, *[L.ptr(a) for a in cal],
                                     ctypes.byref(s), None))
