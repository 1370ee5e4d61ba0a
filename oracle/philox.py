"""Philox4x32-10 counter-based generator and the draw layout of the CRV chain (oracle; test infrastructure only).

Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11): a
128-bit counter (c0..c3) and a 64-bit key (k0, k1), ten rounds of two 32x32->64 multiplies with the multipliers
0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 / 0xBB67AE85. Vectorised over numpy uint32 arrays.

Draw layout (shared with the HIP kernels; every random number of the chain is a pure function of these):
  key    = (seed low 32 bits, kind + 16 * (event key >> 32))
  counter= (event key low 32 bits, channel or FEB key, step ordinal in the event or window tick, sub-index)
The kinds below name what a draw decides. A uniform is u = (x + 0.5) 2^-32 for a 32-bit output x, so u is in
(0, 1) and exact in double.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)

KIND_COUNT = 1     # Poisson photoelectron count per (step, SiPM)
KIND_DELAY = 2     # ordered emission delays per (step, SiPM), four photons per call
KIND_PIXEL = 3     # photon-map bin, two half-pixel bits, fire decision per photon
KIND_XTALK = 4     # four cross-talk proposals of a fired pixel
KIND_XTFIRE = 5    # fire decision of a cross-talk candidate pixel
KIND_TDC0 = 6      # FEB clock phase per (event, ROC, FEB)
KIND_NOISE = 7     # electronics noise, two samples per call
KIND_THERMAL_N = 8  # dark-noise count per SiPM (Poisson, the photon-count sampler)
KIND_THERMAL = 9    # dark discharge k of a SiPM: time, pixel x, pixel y
KIND_TRAP = 10      # trap capture of a firing: trap 0 test and lifetime, trap 1 test and lifetime
KIND_AFTER = 11     # fire decision of a trap afterpulse: trap 0, trap 1
KIND_LCOUNT = 12    # lookup optics: photon number of one type per (step, SiPM) (Poisson or Gauss)
KIND_LBINOM = 13    # lookup optics: Binomial arrivals per substep, one uniform per type and inversion chunk
KIND_LDELAY = 14    # lookup optics: delay of each arriving photon, four per call
DARK_J = 0xFFFFFFFF  # step ordinal of dark discharges in the photon identity (photons have smaller ones)


def philox4x32(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 of uint32 arrays (broadcast); returns four uint32 arrays."""
    c = [np.asarray(v, np.uint64) & MASK for v in np.broadcast_arrays(c0, c1, c2, c3)]
    k0 = np.asarray(k0, np.uint64) & MASK
    k1 = np.asarray(k1, np.uint64) & MASK
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(W0)) & MASK
            k1 = (k1 + np.uint64(W1)) & MASK
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & MASK]
    return tuple(v.astype(np.uint32) for v in c)


def uniform(x):
    """(x + 0.5) * 2^-32 in double: (0, 1), exact."""
    return (np.asarray(x, np.uint32).astype(np.float64) + 0.5) * 2.0 ** -32


def key(seed, kind, evkey=0):
    ev = np.asarray(evkey, np.int64)
    return np.uint64(seed) & MASK, (np.uint64(kind) + (ev >> 32).astype(np.uint64) * np.uint64(16)) & MASK


def draw(seed, kind, evkey, c1, c2, c3):
    """Four uniforms (arrays) of one Philox call per element."""
    ev = np.asarray(evkey, np.int64)
    k0, k1 = key(seed, kind, ev)
    out = philox4x32(ev.astype(np.uint64) & MASK, np.asarray(c1, np.int64).astype(np.uint64) & MASK,
                     np.asarray(c2, np.int64).astype(np.uint64) & MASK,
                     np.asarray(c3, np.int64).astype(np.uint64) & MASK, k0, k1)
    return tuple(uniform(v) for v in out)


def pack_photon(rank, gen=0, pix=0):
    """Sub-index of a photon-level draw: photon rank in its (step, SiPM) list, cascade generation, pixel id."""
    return (np.asarray(rank, np.int64) << 16) | (np.asarray(gen, np.int64) << 11) | np.asarray(pix, np.int64)
