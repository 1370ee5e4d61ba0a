"""Checks of a Gumbel fit's end point against the reference's recorded fit (test helper).

The reference's fit (reco/pulsefit.py:133-216, reco/pulsefit_cpu.py:5-57) runs a fixed 40 Levenberg-Marquardt and
8 Gauss-Newton iterations and accepts a step only if the double-precision chi2 goes down. Near the minimum chi2 is
flat to within its own rounding over a parameter range of a few 1e-9 relative along the fit's softest direction, so
where the reference stops inside that range is set by the last bits of its exp, sin/cos and LAPACK solve. Two end
points are equivalent when their exact chi2 values differ by less than the rounding noise of evaluating chi2 in
double precision, eps * (chi2 + 2 sum |r| |f|) (r the residuals, f the model values).
"""
import numpy as np

EPS = np.finfo(np.float64).eps
LD = np.longdouble
BAR = 1e-9                   # relative agreement asked of pulse PE, height and time


def chi2_and_noise(t, y, mask, A, mu, b):
    """chi2 of the Gumbel A exp(-z - e^-z), z = (t - mu) / b, over the masked points, in extended precision, and the
    rounding noise of its double-precision evaluation (chi2 stays a long double, so differences keep their bits)."""
    t, y, m = np.asarray(t, LD), np.asarray(y, LD), np.asarray(mask, LD)
    z = (t - LD(mu)) / LD(b)
    f = LD(A) * np.exp(-z - np.exp(-z))
    r = (y - f) * m
    chi2 = (r * r).sum()
    return chi2, float(EPS * (chi2 + 2 * (np.abs(r) * np.abs(f) * m).sum()))


def window_map(windows):
    """(waveform row, pulse seq) -> (t, y, mask) of its fit, from the oracle's `windows` list."""
    out = {}
    for rows, seq, tt, yy, mask in windows:
        for i, w in enumerate(rows):
            out[(int(w), int(seq))] = (tt[i], yy[i], mask[i])
    return out


# Largest relative differences between two chi2-equivalent end points seen on the reference's golden batches
# (synth12 sets each of them; the other batches stay below half): PE = A b / area moves 3.1e-9, A and the height
# A / e 2.2e-9, beta = b 5.2e-9. Times agree within 1.5e-11 relative.
BOUNDS = {"pe": 3.5e-9, "height": 2.5e-9, "peph": 2.5e-9, "fit_A": 2.5e-9, "fit_b": 6e-9, "beta": 6e-9}


def check_fit_values(P, ref, windows, bounds=BOUNDS):
    """P and ref: pulse tables in the same (waveform, seq) order with identical flags. Times within 1e-9 relative.
    Each value of `bounds` either agrees to 1e-9 relative or belongs to a pulse whose two end points are
    chi2-equivalent (|chi2(P) - chi2(ref)| <= noise), and stays within its bound. Returns the largest relative
    difference of PE and height and the largest chi2 difference over noise among the pulses that needed the
    equivalence."""
    W = window_map(windows)
    for k in ("time", "letime", "fit_mu"):
        np.testing.assert_allclose(P[k], ref[k], rtol=BAR, atol=0, err_msg=k)
    rel = np.zeros(len(P["wf"]))
    for k, bound in bounds.items():
        a, b = np.asarray(P[k], float), np.asarray(ref[k], float)
        d = np.where(a == b, 0.0, np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        assert d.max(initial=0.0) <= bound, f"{k} differs by {d.max():.3g} relative (bound {bound:g})"
        rel = np.maximum(rel, d)
    worst = 0.0
    for i in np.flatnonzero(rel > BAR):
        t, y, m = W[(int(P["wf"][i]), int(P["seq"][i]))]
        ca, na = chi2_and_noise(t, y, m, P["fit_A"][i], P["fit_mu"][i], P["fit_b"][i])
        cb, nb = chi2_and_noise(t, y, m, ref["fit_A"][i], ref["fit_mu"][i], ref["fit_b"][i])
        ratio = float(abs(ca - cb)) / max(na, nb)
        assert ratio <= 1.0, (f"pulse {i} (waveform {P['wf'][i]}, seq {P['seq'][i]}) differs by {rel[i]:.3g} "
                              f"relative and its chi2 by {ratio:.3g} x the rounding noise")
        worst = max(worst, ratio)
    peh = max((float(np.max(np.abs(np.asarray(P[k]) - ref[k]) / np.abs(ref[k]), initial=0.0))
               for k in ("pe", "height")), default=0.0)
    return peh, worst
