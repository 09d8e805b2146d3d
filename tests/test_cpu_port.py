"""The C restatement used as the CPU baseline (oracle/cpu_port.c) agrees with
the NumPy oracle: same Philox draws, same discrete trajectory over the first
steps, and M-line force within 1 % while the bound counts agree (the port
is FP32 with the reference's 0.75 pN stopping floor, the oracle FP64)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import mf_oracle as mo

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def port():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "..", "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("cannot build oracle/libmfc.so: " + r.stderr[-300:])
    from oracle import cpu_port
    return cpu_port


def test_port_matches_oracle(port, topo22, pdict):
    T, B = 15, 2
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0], drv[..., 2] = 1000.0, 14.0
    drv[0, :, 1], drv[1, :, 1] = 4.5, 5.5
    seeds = np.array([3, 4], np.uint64)
    force, resid, iters, nbound = port.run(topo22, pdict, drv, seeds, threads=2)
    tol = mo.solver_tolerance(pdict)
    assert np.all((resid <= tol * 1.01) | (iters == 4))
    for b in range(B):
        _, out = mo.run(topo22, pdict, drv[b, :, 0], drv[b, :, 1], drv[b, :, 2], 1.0, int(seeds[b]),
                        mo.SolverConfig(mode="exact"))
        same = nbound[b] == out[:, 5]
        n = T if same.all() else int(np.argmin(same))
        assert n >= 8, f"instance {b}: bound counts differ from step {n}"
        assert np.all(np.abs(force[b, :n] - out[:n, 0]) <= 0.01 * np.abs(out[:n, 0]) + 1.0)


def test_port_threads_do_not_change_results(port, topo22, pdict):
    T, B = 6, 4
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = 950.0, 4.8, 14.0
    seeds = np.arange(B, dtype=np.uint64)
    a = port.run(topo22, pdict, drv, seeds, threads=1)
    b = port.run(topo22, pdict, drv, seeds, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
