"""Single-PE waveform template: the synthetic template file and the reference's resampling to 0.5 ns.

The reference reads Offline's singlePEWaveform_v3.txt (two columns: time in ns, voltage in V) and resamples it with
MakeCrvWaveforms::LoadSinglePEWaveform (reco/waveform.py:46-62): time stretched by singlePEWaveformStretchFactor,
linear interpolation onto a `precision` grid, cut at singlePEWaveformMaxTime. That file is not available here, so
the synthetic detector uses a Gumbel-shaped single-PE pulse whose fitted shape matches the nominal calibration
(peak 11.4 ADC = 4.96 mV, area 394.6 ADC ns, i.e. width 12.73 ns after the 1.047 stretch), written in the same
two-column format and resampled by the same rule.
"""
import os

import numpy as np

PEAK_V = 11.4 / 2300.0
RAW_BETA = 394.6 / (11.4 * np.e) / 1.047      # ns, before the stretch
RAW_PEAK_T = 30.0 / 1.047                      # ns, before the stretch


def synthetic_raw(step=0.1, t_max=100.0):
    """(t, v) columns of the synthetic template file."""
    t = np.arange(0.0, t_max + 0.5 * step, step)
    z = (t - RAW_PEAK_T) / RAW_BETA
    return t, PEAK_V * np.exp(1.0 - z - np.exp(-z))


def write_raw(path, step=0.1):
    t, v = synthetic_raw(step)
    np.savetxt(path, np.column_stack([t, v]), fmt="%.6f %.9e")


def resample(raw_t, raw_v, precision=0.5, stretch=1.047, max_time=100.0):
    """Reference reco/waveform.py:47-62 on the two template columns."""
    t = np.asarray(raw_t, np.float64) * stretch
    v = np.asarray(raw_v, np.float64)
    out = [v[0]]
    idx = 1
    for i in range(1, len(t)):
        tt = idx * precision
        while t[i] >= tt and idx * precision < max_time:
            frac = (tt - t[i - 1]) / (t[i] - t[i - 1])
            out.append((v[i] - v[i - 1]) * frac + v[i - 1])
            idx += 1
            tt = idx * precision
        if idx * precision >= max_time:
            break
    return np.asarray(out)


def load(cfg):
    """The resampled template of a WaveformConfig: its file, or the synthetic pulse when single_pe_file is
    'synthetic'."""
    if cfg.single_pe_file == "synthetic" or not os.path.exists(cfg.single_pe_file):
        if cfg.single_pe_file != "synthetic":
            raise FileNotFoundError(cfg.single_pe_file)
        t, v = synthetic_raw()
    else:
        raw = np.loadtxt(cfg.single_pe_file)
        t, v = raw[:, 0], raw[:, 1]
    return resample(t, v, cfg.single_pe_precision, cfg.single_pe_stretch, cfg.single_pe_max_time)
