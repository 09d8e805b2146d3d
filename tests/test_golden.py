"""Committed fixtures (tests/golden/, written by tests/golden/make_fixtures.py):
the measured workloads' condition tables, and an oracle trajectory that pins
the oracle against regressions."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture():
    with open(os.path.join(HERE, "conditions.json")) as f:
        return json.load(f)


def test_bench_draws_the_committed_conditions():
    import bench
    fx = _fixture()
    drv = bench.conditions(256, 1)
    assert np.array_equal(drv[:, 0, 1], np.asarray(fx["sweep"]["pCa"], np.float32))
    assert np.array_equal(drv[:, 0, 0], np.asarray(fx["sweep"]["z_line"], np.float32))
    lat, lengths, pCa = bench.uneven_conditions()
    assert np.array_equal(lat, fx["uneven"]["lattice_index"])
    assert np.array_equal(lengths, fx["uneven"]["lengths"])
    assert np.array_equal(pCa, np.asarray(fx["uneven"]["pCa"], np.float32))
    # uneven: 96 of each lattice, mean length ~600 steps (230k sarcomere-steps)
    assert np.bincount(lat).tolist() == [96] * 4 and 200 <= lengths.min() and lengths.max() <= 1000
    assert np.allclose(np.logspace(0, 3, 64), fx["sinusoid"]["freq_hz"], rtol=0, atol=0)


def test_oracle_reproduces_its_recorded_trajectory(topo22, pdict):
    from oracle import mf_oracle as mo
    g = np.load(os.path.join(HERE, "oracle_2x2_pCa45.npz"))
    want = g["metrics"]
    T = want.shape[0]
    _, out = mo.run(topo22, pdict, np.full(T, float(g["z_line"])), np.full(T, float(g["pCa"])),
                    np.full(T, 14.0), 1.0, int(g["seed"]), mo.SolverConfig(mode="exact"))
    i_counts = [mo.METRIC_NAMES.index(n) for n in ("n_bound", "n_xb_drx", "n_xb_loose", "n_xb_tight_1",
                                                   "n_xb_tight_2", "n_tm_state_3")]
    assert np.array_equal(out[:, i_counts], want[:, i_counts])
    assert np.allclose(out, want, rtol=1e-9, atol=1e-9, equal_nan=True)
    assert want[-1, mo.METRIC_NAMES.index("n_bound")] > 0
