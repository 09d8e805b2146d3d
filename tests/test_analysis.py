"""Complex-stiffness post-processing (examples/sinusoidal_analysis.py:131-212)
on synthetic traces with a known answer."""
import numpy as np

from multifil_jax_amd.analysis import complex_stiffness, sinusoid_traces


def test_recovers_known_complex_stiffness():
    dt, T = 0.1, 2000
    freqs = np.logspace(0, 3, 64)
    z = sinusoid_traces(freqs, 1000.0, 2.5, dt, T)
    t = np.arange(T) * dt
    Y = (3.0 + 2.0j) * (1 + freqs / 500.0)                 # per-frequency target
    w = 2 * np.pi * freqs[:, None] / 1000.0
    # F = Re(Y) (z - mean) + Im(Y) A cos(wt) = |Y| A sin(wt + angle(Y)) + const
    force = 50.0 + Y.real[:, None] * (z - 1000.0) + Y.imag[:, None] * 2.5 * np.cos(w * t)
    # the reference's rule (one cycle of the lowest frequency, 1000 ms at
    # 1 Hz) leaves no window in a 200 ms run
    ref_rule = complex_stiffness(force, freqs, dt, 1000.0, 2.5, settling_cycles=1)
    assert ref_rule['too_short'].all()
    out = complex_stiffness(force, freqs, dt, 1000.0, 2.5, settling_ms=20.0)
    ok = ~out['too_short']
    # 180 ms after settling: fewer than one cycle below 1/0.18 s = 5.6 Hz
    assert not ok[0] and out['n_cycles'][0] == 0 and np.isnan(out['magnitude'][0])
    assert np.array_equal(ok, freqs * 0.18 >= 1.0)
    # a window of whole cycles rounded to the dt grid leaks ~1e-4 at
    # frequencies whose period is not a whole number of samples
    assert np.allclose(out['magnitude'][ok], np.abs(Y[ok]), rtol=2e-3)
    assert np.allclose(out['phase'][ok], np.angle(Y[ok]), atol=2e-3)
    # exact where every cycle is a whole number of samples
    fx = np.array([10.0, 50.0, 100.0, 250.0, 500.0, 1000.0])
    zx = sinusoid_traces(fx, 1000.0, 2.5, dt, T)
    Yx = (3.0 + 2.0j) * (1 + fx / 500.0)
    fxt = 50.0 + Yx.real[:, None] * (zx - 1000.0) + Yx.imag[:, None] * 2.5 * np.cos(2 * np.pi * fx[:, None] / 1000.0 * t)
    ex = complex_stiffness(fxt, fx, dt, 1000.0, 2.5, settling_ms=20.0)
    assert np.allclose(ex['elastic'], Yx.real, rtol=1e-9) and np.allclose(ex['viscous'], Yx.imag, rtol=1e-9)


def test_short_window_high_frequencies_only():
    dt, T = 0.1, 2000
    freqs = np.array([1.0, 5.0, 100.0, 1000.0])
    force = np.zeros((4, T))
    out = complex_stiffness(force, freqs, dt, 1000.0, 2.5, settling_cycles=0)
    # with no settling, 200 ms holds 0 cycles at 1 Hz, 1 at 5 Hz
    assert list(out['n_cycles']) == [0, 1, 20, 200]
    assert list(out['too_short']) == [True, False, False, False]
