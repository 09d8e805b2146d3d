"""The GPU engine against the committed oracle trajectory
(tests/golden/oracle_2x2_pCa45.npz): same Philox draws, refined solver.
Bars: identical discrete state counts at every step; M-line force within
1e-5 relative; lattice spacing and z_line exact."""
import os

import numpy as np
import pytest

from oracle import mf_oracle as mo

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_engine_matches_the_recorded_oracle_trajectory(topo22, skeletal):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    from gpu_helpers import drivers_const
    g = np.load(os.path.join(HERE, "oracle_2x2_pCa45.npz"))
    want = g["metrics"]
    T = want.shape[0]
    _, d = skeletal
    drv = drivers_const(1, T, float(g["z_line"]), float(g["pCa"]))
    bt = Batch(LatticePack([topo22]), np.zeros(1, np.int32), d.rows(1), drv,
               seeds=np.array([int(g["seed"])], np.uint64))
    bt.run(T, cfg=SolverConfig(mode="refined"))
    m, f = bt.metrics()[0], bt.force()[0]
    assert np.array_equal(m[:, 5:12], want[:, 5:12].astype(np.float32))
    assert np.array_equal(m[:, 19:23], want[:, 19:23].astype(np.float32))
    assert np.max(np.abs(f - want[:, 0]) / np.abs(want[:, 0])) < 1e-5
    assert np.array_equal(m[:, 2:5], want[:, 2:5].astype(np.float32))
