"""Lattices above 6x6 on the GPU: 8x8 and 10x10 (the reference's dynamic
spacing benchmark runs 2x2 to 10x10, examples/benchmarks/
benchmark_dynamic_ls.py:136-137) and the larger flight-muscle lattices,
Lethocerus 4x4 and Drosophila 4x6.  Their positions, per-head state and TM
states live in the device workspace (placement level 2), only the
atomically updated site arrays in LDS.

Bars (refined solver, identical Philox draws): while the XB state counts
agree (required for the first 4 steps), the M-line force is within 1e-5
relative of the FP64 oracle and the solved spacing within 1e-6 nm (+ the
float32 rounding of the metric); residual <= 1e-5 pN at every step.
"""
import time

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


def _case(name):
    from multifil_jax_amd.params import (get_skeletal_params, get_lethocerus_params,
                                         get_drosophila_params, DYNAMIC_FIELDS)
    from multifil_jax_amd.topology import SarcTopology
    fn, shape, z = {"8x8": (get_skeletal_params, (8, 8), 1000.0),
                    "10x10": (get_skeletal_params, (10, 10), 1000.0),
                    "lethocerus4x4": (get_lethocerus_params, (4, 4), 1544.0),
                    "drosophila4x6": (get_drosophila_params, (4, 6), 1544.0)}[name]
    s, d = fn()
    topo = SarcTopology.create(*shape, s, d)
    return topo, d, {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}, z


@pytest.mark.parametrize("name,dynamic,T", [("8x8", False, 12), ("8x8", True, 12), ("10x10", False, 6),
                                            ("lethocerus4x4", False, 16), ("drosophila4x6", True, 16)])
def test_huge_lattice_free_run_refined(torch_cuda, name, dynamic, T):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    topo, d, p, z = _case(name)
    drv = drivers_const(1, T, z, 4.5)
    K = 5.0 * topo.n_thick
    print(f"{name} dynamic={dynamic} T={T}", flush=True)
    bt = Batch(LatticePack([topo]), np.zeros(1, np.int32), d.rows(1), drv, seeds=np.array([5], np.uint64),
               dynamic=(K, 0.0) if dynamic else None)
    t0 = time.perf_counter()
    bt.run(T, cfg=SolverConfig(mode="refined"))
    print(f"  device: {(time.perf_counter() - t0) / T * 1e3:.2f} ms per step (incl. first-call setup)", flush=True)
    assert bt.d_ws is not None
    fg, mg = bt.force()[0], bt.metrics()[0]
    _, out = mo.run(topo, p, drv[0, :, 0], drv[0, :, 1], drv[0, :, 2], 1.0, 5,
                    mo.SolverConfig(mode="exact"), K_lat=K if dynamic else None, nu=0.0)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1)
    assert same[:4].all()
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.max(np.abs(mg[same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6
    assert np.all(mg[:, 1] <= 1e-5)
    if name in ("8x8", "10x10"):
        assert out[-1, 5] > 0
