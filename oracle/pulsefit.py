"""CRV pulse reconstruction oracle (test infrastructure only).

Restates the reference's reco/pulsefit.py:49-316 (near_threshold_peak, round_half_away, gumbel, find_next_peak,
_to_ext / _to_int, fit_gumbel, reconstruct, reconstruct_full, filter_channels) and reco/pulsefit_cpu.py:5-57
(fit_cached: 40 Levenberg-Marquardt iterations, 8 Gauss-Newton polish steps with step halving 1, 1/2, 1/4, 1/8,
EDM <= 1e-3 max(chi2, 1), Hessian eigenvalue check). Arithmetic is fixed so that a HIP kernel reproduces it:
sums over fit points run sequentially in window order, the 3x3 systems are solved by Gaussian elimination with
partial pivoting written out (LAPACK getrf's pivot rule), the eigenvalues by six cyclic Jacobi sweeps, and the
transcendentals come from oracle.fmath. A singular system (zero pivot) affects its own waveform only: the LM step
is zero, the polish stops, the EDM is infinite (the reference raises for the whole batch there).
"""
from dataclasses import dataclass, replace

import numpy as np

from . import fmath as F

PERIOD = 12.5
E = np.e


@dataclass(frozen=True)
class RecoConfig:
    min_adc_difference: float = 40.0
    default_beta: float = 19.0
    min_beta: float = 5.0
    max_beta: float = 40.0
    max_time_difference: float = 15.0
    min_pulse_height_ratio: float = 0.7
    max_pulse_height_ratio: float = 1.5
    le_time_factor: float = 0.985
    samples_before: int = 4
    samples_after: int = 3
    pedestal_undershoot_threshold: float = 6.0
    use_time_offset_db: bool = True
    time_offset_scale: float = 1.0
    ignore_channels: bool = True
    max_pulses: int = 8
    lm_iterations: int = 40
    limit_tolerance: float = 0.01
    knife_edge_adc: float = 1.0


def near_threshold_peak(y, valid, thr, edge):
    yv = np.where(valid, y, -1e9)
    left = np.concatenate([np.full((y.shape[0], 1), -1e9), yv[:, :-1]], 1)
    right = np.concatenate([yv[:, 1:], np.full((y.shape[0], 1), -1e9)], 1)
    peak = valid & (yv >= left) & (yv > right)
    return (peak & (np.abs(yv - thr) <= edge)).any(1)


def round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def gumbel(t, A, mu, b):
    z = (t - mu) / b
    return A * F.exp(-z - F.exp(-z))


def find_next_peak(y, n, start, active, cfg):
    W, L = y.shape
    rows = np.arange(W)
    found = np.zeros(W, bool)
    rising = np.zeros(W, bool)
    done = ~active
    ps = np.zeros(W, np.int64)
    pe = np.zeros(W, np.int64)
    for b in range(1, L):
        live = ~done & (b > start) & (b < n)
        if not live.any():
            continue
        prev, cur = y[:, b - 1], y[:, b]
        up = live & (prev < cur)
        rising |= up
        ps[up] = b
        pe[up] = b
        flat = live & (prev == cur) & rising
        pe[flat] = b
        down = live & (prev > cur) & rising
        hit = down & (y[rows, pe] > cfg.min_adc_difference)
        found |= hit
        done |= hit
        rising &= ~(down & ~hit)
    fs = ps - 1
    fe = pe + 1
    go = found.copy()
    for k in range(1, cfg.samples_before + 1):
        i = ps - k
        ok = go & (i >= 0) & (i >= ps - cfg.samples_before)
        ii = np.clip(i, 0, L - 2)
        cond = ok & (y[rows, ii] <= y[rows, ii + 1])
        fs[cond] = i[cond]
        go = cond & (i != 0)
    go = found.copy()
    for k in range(1, cfg.samples_after + 1):
        i = pe + k
        ok = go & (i < n)
        ii = np.clip(i, 1, L - 1)
        cond = ok & (y[rows, ii - 1] >= y[rows, ii])
        fe[cond] = i[cond]
        go = cond
    fs = np.where(found & (ps - fs > 1), fs + 1, fs)
    fe = np.where(found & (fe - pe > 1), fe - 1, fe)
    return found, ps, pe, fs, fe


def _to_ext(u, lo, hi):
    return lo + (hi - lo) * (F.sin(u) + 1.0) * 0.5


def _to_int(p, lo, hi):
    return F.asin(np.clip(2.0 * (p - lo) / (hi - lo) - 1.0, -1.0, 1.0))


def solve3(A, b):
    """x with A x = b for [K,3,3] / [K,3]; Gaussian elimination, partial pivoting (first max |a|), written out.
    Returns (x, singular)."""
    a = np.array(A, np.float64, copy=True)
    x = np.array(b, np.float64, copy=True)
    K = a.shape[0]
    rows = np.arange(K)
    sing = np.zeros(K, bool)
    for c in range(3):
        piv = np.full(K, c)
        best = np.abs(a[:, c, c])
        for r in range(c + 1, 3):
            m = np.abs(a[:, r, c]) > best
            piv = np.where(m, r, piv)
            best = np.where(m, np.abs(a[:, r, c]), best)
        sw = piv != c
        if sw.any():
            ra, xa = a[rows, c].copy(), x[rows, c].copy()
            a[rows, c], x[rows, c] = a[rows, piv], x[rows, piv]
            a[rows, piv], x[rows, piv] = ra, xa
        sing |= a[:, c, c] == 0.0
        dd = np.where(a[:, c, c] == 0.0, 1.0, a[:, c, c])
        for r in range(c + 1, 3):
            l = a[:, r, c] / dd
            for cc in range(c + 1, 3):
                a[:, r, cc] = a[:, r, cc] - l * a[:, c, cc]
            x[:, r] = x[:, r] - l * x[:, c]
    d2 = np.where(sing, 1.0, a[:, 2, 2])
    d1 = np.where(sing, 1.0, a[:, 1, 1])
    d0 = np.where(sing, 1.0, a[:, 0, 0])
    x2 = x[:, 2] / d2
    x1 = (x[:, 1] - a[:, 1, 2] * x2) / d1
    x0 = ((x[:, 0] - a[:, 0, 1] * x1) - a[:, 0, 2] * x2) / d0
    return np.stack([x0, x1, x2], 1), sing


def eig3_minmax(S):
    """(min, max) eigenvalue of symmetric [K,3,3] by six cyclic Jacobi sweeps over (0,1), (0,2), (1,2)."""
    a = np.array(S, np.float64, copy=True)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        return _jacobi(a)


def _jacobi(a):
    for _ in range(6):
        for p, q in ((0, 1), (0, 2), (1, 2)):
            apq = a[:, p, q]
            nz = apq != 0.0
            den = np.where(nz, 2.0 * apq, 1.0)
            th = (a[:, q, q] - a[:, p, p]) / den
            t = np.where(th >= 0.0, 1.0, -1.0) / (np.abs(th) + np.sqrt(th * th + 1.0))
            t = np.where(nz, t, 0.0)
            c = 1.0 / np.sqrt(t * t + 1.0)
            s = t * c
            app = a[:, p, p] - t * apq
            aqq = a[:, q, q] + t * apq
            r = 3 - p - q
            arp, arq = a[:, r, p], a[:, r, q]
            nrp = c * arp - s * arq
            nrq = s * arp + c * arq
            a[:, p, p], a[:, q, q] = app, aqq
            a[:, p, q] = a[:, q, p] = 0.0
            a[:, r, p] = a[:, p, r] = nrp
            a[:, r, q] = a[:, q, r] = nrq
    d = np.stack([a[:, 0, 0], a[:, 1, 1], a[:, 2, 2]], 1)
    return d.min(1), d.max(1)


def _sums(J, r):
    """chi2, JTJ [K,3,3], JTr [K,3] with sequential sums over the points."""
    K, P = r.shape
    chi2 = np.zeros(K)
    JTJ = np.zeros((K, 3, 3))
    JTr = np.zeros((K, 3))
    for p in range(P):
        chi2 = chi2 + r[:, p] * r[:, p]
        for i in range(3):
            JTr[:, i] = JTr[:, i] + J[:, p, i] * r[:, p]
            for j in range(i, 3):
                JTJ[:, i, j] = JTJ[:, i, j] + J[:, p, i] * J[:, p, j]
    for i in range(3):
        for j in range(i):
            JTJ[:, i, j] = JTJ[:, j, i]
    return chi2, JTJ, JTr


def fit_gumbel(t, y, mask, p0, lo, hi, iterations=40):
    """Least-squares Gumbel fits. t, y, mask [K,P]; p0, lo, hi [K,3]. Returns params, chi2, converged, hess_ok."""
    K = t.shape[0]
    u = _to_int(p0, lo, hi)
    lam = np.full(K, 1e-3)
    w = mask.astype(float)

    def resid_jac(u):
        p = _to_ext(u, lo, hi)
        A, mu, b = p[:, 0:1], p[:, 1:2], p[:, 2:3]
        z = (t - mu) / b
        ez = F.exp(-z)
        g = F.exp(-z - ez)
        f = A * g
        r = (y - f) * w
        dfdz = f * (ez - 1.0)
        dfdmu = dfdz * (-1.0 / b)
        dfdb = dfdz * (-z / b)
        dpdu = (hi - lo) * 0.5 * F.cos(u)
        J = np.stack([g * dpdu[:, 0:1], dfdmu * dpdu[:, 1:2], dfdb * dpdu[:, 2:3]], axis=-1) * w[..., None]
        return r, J

    r, J = resid_jac(u)
    chi2, JTJ, JTr = _sums(J, r)
    eye = np.eye(3)
    for _ in range(iterations):
        D = np.stack([JTJ[:, 0, 0], JTJ[:, 1, 1], JTJ[:, 2, 2]], 1)
        Aug = JTJ.copy()
        for i in range(3):
            Aug[:, i, i] = JTJ[:, i, i] + lam * (D[:, i] + 1e-12)
        du, sing = solve3(Aug, JTr)
        du[sing] = 0.0
        un = u + du
        rn, Jn = resid_jac(un)
        chi2n, JTJn, JTrn = _sums(Jn, rn)
        better = chi2n < chi2
        u = np.where(better[:, None], un, u)
        r = np.where(better[:, None], rn, r)
        J = np.where(better[:, None, None], Jn, J)
        chi2 = np.where(better, chi2n, chi2)
        JTJ = np.where(better[:, None, None], JTJn, JTJ)
        JTr = np.where(better[:, None], JTrn, JTr)
        lam = np.clip(np.where(better, lam * 0.3, lam * 10.0), 1e-9, 1e9)
    stopped = np.zeros(K, bool)
    for _ in range(8):
        Ag = JTJ + 1e-12 * eye
        du, sing = solve3(Ag, JTr)
        stopped |= sing
        du = np.where(np.isfinite(du) & ~stopped[:, None], du, 0.0)
        for scale in (1.0, 0.5, 0.25, 0.125):
            un = u + scale * du
            rn, Jn = resid_jac(un)
            chi2n, JTJn, JTrn = _sums(Jn, rn)
            better = (chi2n < chi2) & ~stopped
            u = np.where(better[:, None], un, u)
            r = np.where(better[:, None], rn, r)
            J = np.where(better[:, None, None], Jn, J)
            chi2 = np.where(better, chi2n, chi2)
            du = np.where(better[:, None], 0.0, du)
            JTJ = np.where(better[:, None, None], JTJn, JTJ)
            JTr = np.where(better[:, None], JTrn, JTr)
    step, sing = solve3(JTJ + 1e-12 * eye, JTr)
    edm = 0.5 * ((JTr[:, 0] * step[:, 0] + JTr[:, 1] * step[:, 1]) + JTr[:, 2] * step[:, 2])
    emin, emax = eig3_minmax(JTJ)
    hess_ok = emin > 1e-12 * np.maximum(emax, 1e-300)
    edm = np.where(sing, np.inf, edm)
    hess_ok &= ~sing
    converged = np.isfinite(edm) & (edm <= 1e-3 * np.maximum(chi2, 1.0))
    return _to_ext(u, lo, hi), chi2, converged, hess_ok


PULSE_KEYS = ("wf", "seq", "pe", "peph", "time", "height", "beta", "chi2", "letime", "failed", "zerondf", "ped",
              "pedfromdb", "fit_A", "fit_mu", "fit_b")


def reconstruct(adc, n, start_tdc, pedestal_db, calib_area, calib_height, cfg=RecoConfig(), windows=None):
    """Reference reco/pulsefit.py:236-316. `windows`, if a list, receives each round's fit inputs
    (pulse rows, t, y, mask), in the order pulses are produced, for checks of the fit's end point."""
    adc = np.asarray(adc)
    W = adc.shape[0]
    L = adc.shape[1] if adc.ndim == 2 else 0
    if W == 0:
        out = {k: np.zeros(0) for k in ('pe', 'peph', 'time', 'height', 'beta', 'chi2', 'letime', 'ped',
                                       'fit_A', 'fit_mu', 'fit_b')}
        out.update({k: np.zeros(0, np.int64) for k in ('wf', 'seq')})
        out.update({k: np.zeros(0, bool) for k in ('failed', 'zerondf', 'pedfromdb', 'wf_near_threshold')})
        return out
    n = np.asarray(n, np.int64)
    start_tdc = np.asarray(start_tdc, np.int64)
    adc = adc.astype(float)
    ped = np.where(pedestal_db - adc[:, 0] > cfg.pedestal_undershoot_threshold, adc[:, 0], pedestal_db)
    ped_from_db = ~(pedestal_db - adc[:, 0] > cfg.pedestal_undershoot_threshold)
    y = adc - ped[:, None]
    y[np.arange(L)[None, :] >= n[:, None]] = 0.0
    tgrid = (start_tdc[:, None] + np.arange(L)[None, :]) * PERIOD
    start = np.zeros(W, np.int64)
    active = np.ones(W, bool)
    failed_before = np.zeros(W, bool)
    valid = np.arange(L)[None, :] < n[:, None]
    near_thr = near_threshold_peak(y, valid, cfg.min_adc_difference, cfg.knife_edge_adc)
    out = {k: [] for k in PULSE_KEYS}
    for seq in range(cfg.max_pulses):
        found, ps, pe, fs, fe = find_next_peak(y, n, start, active, cfg)
        if not found.any():
            break
        k = np.flatnonzero(found)
        P = int((fe[k] - fs[k]).max()) + 1
        yps = y[k, ps[k]]
        t_ps = (start_tdc[k] + ps[k]) * PERIOD
        t_pe = (start_tdc[k] + pe[k]) * PERIOD
        t_peak = 0.5 * (t_ps + t_pe)
        p0 = np.stack([yps * E, t_peak, np.full(len(k), cfg.default_beta)], 1)
        lo = np.stack([yps * E * cfg.min_pulse_height_ratio, t_ps - cfg.max_time_difference,
                       np.full(len(k), cfg.min_beta)], 1)
        hi = np.stack([yps * E * cfg.max_pulse_height_ratio, t_pe + cfg.max_time_difference,
                       np.full(len(k), cfg.max_beta)], 1)
        idx = fs[k, None] + np.arange(P)[None, :]
        mask = idx <= fe[k, None]
        idxc = np.clip(idx, 0, L - 1)
        tt = tgrid[k[:, None], idxc]
        yy = y[k[:, None], idxc]
        par, chi2, conv, hok = fit_gumbel(tt, yy, mask, p0, lo, hi, cfg.lm_iterations)
        if windows is not None:
            windows.append((k, seq, tt, yy, mask))
        npts = mask.sum(1)
        ndf = npts - 3
        span = hi - lo
        near = ((par - lo) / span < cfg.limit_tolerance) | ((hi - par) / span < cfg.limit_tolerance)
        failed = ~conv | ~hok | near.any(1)
        A, mu, b = par[:, 0], par[:, 1], par[:, 2]
        pe_ = np.where(failed, yps * E * cfg.default_beta, A * b) / calib_area[k]
        tpulse = np.where(failed, t_peak, mu)
        height = np.where(failed, yps, A / E)
        beta = np.where(failed, cfg.default_beta, b)
        c2 = np.where(failed | (ndf <= 0), -1.0, chi2 / np.maximum(ndf, 1))
        failed_seq = failed | failed_before[k]
        for key, v in (("wf", k), ("seq", np.full(len(k), seq)), ("pe", pe_), ("peph", height / calib_height[k]),
                       ("time", tpulse), ("height", height), ("beta", beta), ("chi2", c2),
                       ("letime", tpulse - cfg.le_time_factor * beta), ("failed", failed_seq), ("zerondf", ndf <= 0),
                       ("ped", ped[k]), ("pedfromdb", ped_from_db[k]), ("fit_A", A), ("fit_mu", mu), ("fit_b", b)):
            out[key].append(v)
        failed_before[k] = failed_seq
        sub = round_half_away(gumbel(tgrid[k], A[:, None], mu[:, None], b[:, None]))
        sub[np.arange(L)[None, :] >= n[k, None]] = 0.0
        y[k] -= sub
        near_thr[k] |= near_threshold_peak(y[k], valid[k], cfg.min_adc_difference, cfg.knife_edge_adc)
        start[k] = fe[k]
        active = found
    bool_keys = ('failed', 'zerondf', 'pedfromdb')
    res = {key: (np.concatenate(v) if v else np.zeros(0, bool if key in bool_keys else
                                                     np.int64 if key in ('wf', 'seq') else float))
           for key, v in out.items()}
    order = np.lexsort((res["seq"], res["wf"])) if len(res["wf"]) else np.zeros(0, np.int64)
    res = {key: v[order] for key, v in res.items()}
    res["wf_near_threshold"] = near_thr
    return res


def reconstruct_full(adc, n, start_tdc, pedestal_db, calib_area, calib_height, cfg=RecoConfig(), windows=None):
    return reconstruct(adc, n, start_tdc, pedestal_db, calib_area, calib_height,
                       replace(cfg, max_pulses=max(cfg.max_pulses, int(np.max(n, initial=0)))), windows)


def filter_channels(pulses, channels, status, cfg=RecoConfig()):
    if not cfg.ignore_channels:
        return pulses
    keep = (status[channels[pulses['wf']]] & 0b010) == 0
    return {k: v if k == 'wf_near_threshold' else v[keep] for k, v in pulses.items()}
