"""6x6 lattice (N 8352, X 5616, S 6480): too large for LDS, so the engine
keeps its per-XB / per-site arrays in a device workspace.  Parity against
the FP64 oracle with identical Philox draws, fixed and dynamic spacing.

Bars: refined solver, while the XB state counts agree, M-line force within
1e-5 relative and the solved spacing within 1e-6 nm (+ float32 rounding of
the metric); residual <= 1e-5 pN at every step.
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo

from gpu_helpers import drivers_const

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


@pytest.fixture(scope="module")
def topo66(skeletal):
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    return SarcTopology.create(6, 6, s, d)


@pytest.mark.parametrize("dynamic", [False, True])
def test_6x6_free_run_refined(torch_cuda, topo66, skeletal, pdict, dynamic):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    _, d = skeletal
    T = 24   # long enough to reach full activation, where the radial sum is largest
    drv = drivers_const(1, T, 1000.0, 4.5)
    K = 5.0 * topo66.n_thick
    bt = Batch(LatticePack([topo66]), np.zeros(1, np.int32), d.rows(1), drv, seeds=np.array([5], np.uint64),
               dynamic=(K, 0.0) if dynamic else None)
    bt.run(T, cfg=SolverConfig(mode="refined"))
    assert bt.d_ws is not None          # the workspace variant ran
    fg, mg = bt.force()[0], bt.metrics()[0]
    _, out = mo.run(topo66, pdict, drv[0, :, 0], drv[0, :, 1], drv[0, :, 2], 1.0, 5,
                    mo.SolverConfig(mode="exact"), K_lat=K if dynamic else None, nu=0.0)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1)
    assert same[:4].all()
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.max(np.abs(mg[same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6
    assert np.all(mg[:, 1] <= 1e-5)
