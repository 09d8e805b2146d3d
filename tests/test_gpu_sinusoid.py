"""Sinusoidal length perturbation (examples/sinusoidal_analysis.py): complex
stiffness from the device's force traces vs the FP64 oracle's, and the full
4x4 x 64-frequency workload's invariants.

Bars: at reduced size (2x2, 500 and 1000 Hz, 300 steps at dt 0.1 ms) the
force traces agree within 1e-5 relative while the XB state counts agree,
and, when they agree throughout, the complex stiffness within 1e-4
relative; at full size (4x4, 64 frequencies log-spaced 1-1000 Hz, 2000
steps) every residual is <= 1e-5 pN and every frequency with a whole cycle
after 20 ms of settling has a finite stiffness.
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


def _run(topo, d, freqs, T, dt, seeds):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    from multifil_jax_amd.analysis import sinusoid_traces
    B = len(freqs)
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0] = sinusoid_traces(freqs, 1000.0, 2.5, dt, T)
    drv[..., 1], drv[..., 2] = 4.5, 14.0
    bt = Batch(LatticePack([topo]), np.zeros(B, np.int32), d.rows(B), drv, seeds=np.asarray(seeds, np.uint64))
    bt.run(T, cfg=SolverConfig(mode="refined", dt=dt))
    return bt, drv


def test_stiffness_matches_oracle(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd.analysis import complex_stiffness
    _, d = skeletal
    freqs, T, dt = np.array([500.0, 1000.0]), 300, 0.1
    bt, drv = _run(topo22, d, freqs, T, dt, [0, 1])
    fg, mg = bt.force(), bt.metrics()
    for b in range(2):
        _, out = mo.run(topo22, pdict, drv[b, :, 0], drv[b, :, 1], drv[b, :, 2], dt, b,
                        mo.SolverConfig(mode="exact"))
        same = np.all(mg[b, :, 6:12] == out[:, 6:12], axis=1)
        n = T if same.all() else int(np.argmin(same))
        assert n >= 50
        assert np.max(np.abs(fg[b, :n] - out[:n, 0]) / np.abs(out[:n, 0])) < 1e-5
        if n == T:
            a = complex_stiffness(fg[b:b + 1, None], freqs[b:b + 1], dt, 1000.0, 2.5, settling_ms=5.0)
            o = complex_stiffness(out[None, None, :, 0], freqs[b:b + 1], dt, 1000.0, 2.5, settling_ms=5.0)
            Ya = a['elastic'] + 1j * a['viscous']
            Yo = o['elastic'] + 1j * o['viscous']
            assert np.all(np.abs(Ya - Yo) <= 1e-4 * np.abs(Yo))


def test_full_size_sinusoid(torch_cuda, skeletal):
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.analysis import complex_stiffness
    s, d = skeletal
    topo = SarcTopology.create(4, 4, s, d)
    freqs, T, dt = np.logspace(0, 3, 64), 2000, 0.1
    bt, _ = _run(topo, d, freqs, T, dt, np.arange(64))
    m = bt.metrics()
    assert np.all(m[:, :, 1] <= 1e-5)
    st = complex_stiffness(bt.force()[:, None, :], freqs, dt, 1000.0, 2.5, settling_ms=20.0)
    ok = ~st['too_short']
    assert np.array_equal(ok, freqs * 0.18 >= 1.0)
    assert np.all(np.isfinite(st['magnitude'][ok]))
