"""Post-processing of sinusoidal length-perturbation runs: complex stiffness
Y(f) = F(f) / L(f) at the stimulus frequency
(examples/sinusoidal_analysis.py:131-212).

After the settling period, each frequency uses the largest whole number of
its cycles that fits the force trace, so the FFT window is periodic (no
leakage). Both traces are mean-removed before the FFT. Y is taken at the
FFT bin nearest the stimulus frequency.

A frequency whose window holds less than one whole cycle after settling
has no estimate: its Y is NaN and ``too_short`` is set. The reference would
index an empty FFT there.
"""
from __future__ import annotations

import numpy as np


def sinusoid_traces(frequencies, mean_z: float, amplitude_nm: float, dt_ms: float, n_steps: int):
    """z_line(t) = mean_z + A sin(2 pi f t) per frequency, t = k dt (ms)."""
    t = np.arange(n_steps) * dt_ms
    return np.stack([mean_z + amplitude_nm * np.sin(2 * np.pi * f / 1000.0 * t) for f in frequencies])


def complex_stiffness(force, frequencies, dt_ms: float, mean_z: float, amplitude_nm: float,
                      settling_cycles: float = 1.0, settling_ms=None):
    """force: (n_freq, replicates, time) or (n_freq, time).  Returns a dict
    with freq_hz, magnitude, magnitude_std, phase, phase_std, elastic,
    viscous, n_cycles and too_short (one entry per frequency).

    The settling period is ``settling_cycles`` cycles of the lowest frequency
    (the reference's rule), or ``settling_ms`` when given.  With 1 Hz in the
    set, the reference's rule discards 1000 ms; a shorter run then has no
    window at any frequency."""
    force = np.asarray(force, dtype=np.float64)
    if force.ndim == 2:
        force = force[:, None, :]
    freqs_in = np.asarray(frequencies, dtype=np.float64)
    if settling_ms is not None:
        settle = int(settling_ms / dt_ms)
    else:
        settle = int(settling_cycles * 1000.0 / freqs_in.min() / dt_ms)
    f_all = force[..., settle:]
    n_samples = f_all.shape[-1]
    mags, phs, ncyc, short = [], [], [], []
    for i, fr in enumerate(freqs_in):
        n_complete = int(n_samples * dt_ms / 1000.0 * fr)
        n_use = min(int(round(n_complete * 1000.0 / fr / dt_ms)), n_samples)
        ncyc.append(n_complete)
        if n_complete < 1 or n_use < 2:
            short.append(True)
            mags.append(np.full(force.shape[1], np.nan))
            phs.append(np.full(force.shape[1], np.nan))
            continue
        short.append(False)
        fd = f_all[i, :, :n_use]
        t = np.arange(n_use) * dt_ms + settle * dt_ms
        zd = mean_z + amplitude_nm * np.sin(2 * np.pi * fr / 1000.0 * t)
        F = np.fft.fft(fd - fd.mean(axis=-1, keepdims=True), axis=-1)
        Z = np.fft.fft(zd - zd.mean())
        bins = np.fft.fftfreq(n_use, dt_ms / 1000.0)
        k = int(np.argmin(np.abs(bins - fr)))
        Y = F[:, k] / Z[k]
        mags.append(np.abs(Y))
        phs.append(np.angle(Y))
    mag = np.array([m.mean() for m in mags])
    ph = np.array([p.mean() for p in phs])
    return {
        'freq_hz': freqs_in, 'magnitude': mag, 'magnitude_std': np.array([m.std() for m in mags]),
        'phase': ph, 'phase_std': np.array([p.std() for p in phs]),
        'elastic': mag * np.cos(ph), 'viscous': mag * np.sin(ph),
        'n_cycles': np.array(ncyc), 'too_short': np.array(short),
    }
