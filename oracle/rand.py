"""Samplers on Philox uniforms with fixed arithmetic (oracle; test infrastructure only).

Every sampler is written as the same IEEE-754 double operation sequence the HIP kernels use (fmath functions,
no contraction), so a count, a delay or a noise sample is a pure function of (seed, draw layout) on both sides.

* Poisson(mu), mu < 10: inversion by sequential search with one uniform (p0 = e^-mu, p_k = p_{k-1} mu / k).
* Poisson(mu), mu >= 10: PTRS, the transformed rejection with squeeze of W. Hoermann, "The transformed rejection
  method for generating Poisson random variables", Insurance: Mathematics and Economics 12 (1993) 39-45; two
  uniforms per trial, two trials per Philox call; log k! from the Stirling series (terms to x^-11) for k + 1 >= 10
  and from the exact factorial below.
* Ordered exponential delays: the Renyi (1953) representation of the order statistics of N i.i.d. Exp(tau),
  S_r = sum_{i<=r} Z_i / (N - i + 1), Z_i = -log u_i, so the delays of one (step, SiPM) come out sorted.
* Normal: Box-Muller, z = sqrt(-2 log u1) cos(2 pi u2).
"""
from fractions import Fraction
import math

import numpy as np

from . import fmath as F
from . import philox as PX

LN_SQRT_2PI = 0.9189385332046728
STIRLING = [float(Fraction(1, 12)), float(Fraction(-1, 360)), float(Fraction(1, 1260)), float(Fraction(-1, 1680)),
            float(Fraction(1, 1188)), float(Fraction(-691, 360360))]
FACT = [float(math.factorial(k)) for k in range(10)]
POISSON_INV_MAX = 10.0
INV_KMAX = 200          # inversion search bound (mu < 10: P(k > 200) is 0 in double)
PTRS_TRIALS = 64        # rejection bound (acceptance > 0.85 per trial); reaching it is reported, never silent


def loggam_int(x):
    """log Gamma(x) for integer-valued x >= 1 (array)."""
    x = np.asarray(x, np.float64)
    small = x < 10.0
    xi = np.clip(x, 1, 9).astype(np.int64) - 1
    ls = F.log(np.asarray(FACT)[xi])
    xs = np.where(small, 10.0, x)
    z = 1.0 / xs
    z2 = z * z
    p = np.full(np.shape(z), STIRLING[-1])
    for c in STIRLING[-2::-1]:
        p = p * z2 + c
    s = z * p
    t1 = (xs - 0.5) * F.log(xs)
    lg = ((t1 - xs) + LN_SQRT_2PI) + s
    return np.where(small, ls, lg)


def poisson(mu, uniforms):
    """Poisson draws. mu [n]; uniforms(call) -> (u0, u1, u2, u3) arrays [n] of Philox call `call` of each draw.

    Returns (k int64 [n], capped bool [n]); capped marks a PTRS draw that hit PTRS_TRIALS (then k = rint(mu)).
    """
    mu = np.asarray(mu, np.float64)
    n = len(mu)
    k = np.zeros(n, np.int64)
    capped = np.zeros(n, bool)
    if n == 0:
        return k, capped
    inv = mu < POISSON_INV_MAX
    # inversion
    ii = np.flatnonzero(inv)
    if len(ii):
        u = uniforms(0)[0][ii]
        m = mu[ii]
        p = F.exp(-m)
        cdf = p.copy()
        kk = np.zeros(len(ii), np.int64)
        go = u > cdf
        while go.any():
            g = np.flatnonzero(go)
            kk[g] += 1
            p[g] = (p[g] * m[g]) / kk[g].astype(np.float64)
            cdf[g] = cdf[g] + p[g]
            go = (u > cdf) & (kk < INV_KMAX)
        k[ii] = kk
    # PTRS
    jj = np.flatnonzero(~inv)
    if len(jj):
        m = mu[jj]
        slam = np.sqrt(m)
        loglam = F.log(m)
        b = 0.931 + 2.53 * slam
        a = -0.059 + 0.02483 * b
        invalpha = 1.1239 + 1.1328 / (b - 3.4)
        vr = 0.9277 - 3.6224 / (b - 2.0)
        out = np.zeros(len(jj), np.int64)
        done = np.zeros(len(jj), bool)
        cache = {}
        for trial in range(PTRS_TRIALS):
            call = trial >> 1
            if call not in cache:
                cache = {call: uniforms(call)}
            uu = cache[call]
            g = np.flatnonzero(~done)
            if not len(g):
                break
            U = uu[2 * (trial & 1)][jj[g]] - 0.5
            V = uu[2 * (trial & 1) + 1][jj[g]]
            us = 0.5 - np.abs(U)
            kf = np.floor((((2.0 * a[g]) / us + b[g]) * U + m[g]) + 0.43)
            acc = (us >= 0.07) & (V <= vr[g])
            rej = (kf < 0) | ((us < 0.013) & (V > us))
            test = ~acc & ~rej
            if test.any():
                t = np.flatnonzero(test)
                gt = g[t]
                lhs = F.log((V[t] * invalpha[gt]) / (a[gt] / (us[t] * us[t]) + b[gt]))
                rhs = (-m[gt] + kf[t] * loglam[gt]) - loggam_int(kf[t] + 1.0)
                acc[t] = lhs <= rhs
            out[g[acc]] = kf[acc].astype(np.int64)
            done[g[acc]] = True
        if not done.all():
            bad = np.flatnonzero(~done)
            out[bad] = np.rint(m[bad]).astype(np.int64)
            capped[jj[bad]] = True
        k[jj] = out
    return k, capped


def ordered_delays(n, tau, uniforms_of):
    """Sorted Exp(tau) delays for one group of n photons (scalar n); uniforms_of(call) -> 4 uniforms (scalars)."""
    out = np.empty(n)
    s = 0.0
    cache = None
    for r in range(n):
        if r & 3 == 0:
            cache = uniforms_of(r >> 2)
        u = cache[r & 3]
        z = -F.log(np.float64(u))
        s = s + z / np.float64(n - r)
        out[r] = tau * s
    return out


def normal(u1, u2):
    return np.sqrt(-2.0 * F.log(u1)) * F.cos(F.TWO_PI * u2)


BINOM_CHUNK_MEAN = 16.0  # largest -size * log(1 - q) of one inversion chunk: P(0) >= e^-16, no underflow
BINOM_KMAX = 1024        # inversion search bound per chunk (the chunk mean is <= 16: never reached in double)
BINOM_MAX_CHUNKS = 1 << 16  # chunks per draw the draw layout holds (a mean of up to ~1M arrivals)


def binomial_chunk(n, q):
    """Chunk size of the Binomial(n, q) inversion, q <= 1/2: the largest m with m * -log(1 - q) <= 16 (at least 1,
    at most n)."""
    lq = F.log(1.0 - q)
    lim = BINOM_CHUNK_MEAN / -lq
    return np.where(lim >= n.astype(np.float64), n, np.maximum(np.floor(lim), 1.0).astype(np.int64)), lq


def binomial(n, p, uniform_of):
    """Binomial(n, p) draws (arrays; p in [0, 1]). uniform_of(chunk) -> uniform array [len(n)] of each draw's chunk.

    Inversion by sequential search (P(0) = (1 - q)^m, P(k + 1) = P(k) * q / (1 - q) * (m - k) / (k + 1)) on chunks of
    m trials whose mean is at most 16, summed over the chunks of n (a sum of Binomial(m_i, q) is Binomial(sum m_i,
    q)); q = min(p, 1 - p) and k -> n - k when p > 1/2. Returns (k int64, ok bool): ok is False when a draw needs
    more chunks than the draw layout holds."""
    n = np.asarray(n, np.int64)
    p = np.asarray(p, np.float64)
    flip = p > 0.5
    q = np.where(flip, 1.0 - p, p)
    k = np.zeros(len(n), np.int64)
    ok = np.ones(len(n), bool)
    act = np.flatnonzero((n > 0) & (q > 0.0))
    if len(act):
        na, qa = n[act], q[act]
        m, lq = binomial_chunk(na, qa)
        nch = (na + m - 1) // m
        ok[act] = nch <= BINOM_MAX_CHUNKS
        r = qa / (1.0 - qa)
        tot = np.zeros(len(act), np.int64)
        for c in range(int(nch.max())):
            g = np.flatnonzero(nch > c)
            size = np.minimum(m[g], na[g] - c * m[g])
            u = uniform_of(c)[act[g]]
            f = F.exp(size.astype(np.float64) * lq[g])
            cdf = f.copy()
            kk = np.zeros(len(g), np.int64)
            go = (u > cdf) & (kk < size)
            while go.any():
                h = np.flatnonzero(go)
                f[h] = ((f[h] * r[g[h]]) * (size[h] - kk[h]).astype(np.float64)) / (kk[h] + 1).astype(np.float64)
                kk[h] += 1
                cdf[h] = cdf[h] + f[h]
                go = (u > cdf) & (kk < size) & (kk < BINOM_KMAX)
            tot[g] += kk
        k[act] = np.where(flip[act], na - tot, tot)
    k = np.where(flip & (q == 0.0), n, k)  # p == 1: every trial
    return k, ok
