"""Invertebrate flight-muscle lattices on the GPU (sarc_geometry.py:915-1005):
Lethocerus 2x2 (4 heads per crown, 2 thin faces per thick, 80 candidate
sites per head, 160 sites per thin strand) and Drosophila 2x3 (three-class
myosin superlattice, sarc_geometry.py:1118-1131).  Parity against the FP64
oracle with identical Philox draws, fixed and dynamic spacing.

Bars (refined solver): while the XB state counts agree, M-line force within
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


def _preset(name):
    from multifil_jax_amd.params import get_lethocerus_params, get_drosophila_params, DYNAMIC_FIELDS
    from multifil_jax_amd.topology import SarcTopology
    fn, shape = {"lethocerus": (get_lethocerus_params, (2, 2)),
                 "drosophila": (get_drosophila_params, (2, 3))}[name]
    s, d = fn()
    topo = SarcTopology.create(*shape, s, d)
    return topo, d, {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}


@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("name", ["lethocerus", "drosophila"])
def test_invertebrate_free_run_refined(torch_cuda, name, dynamic):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    topo, d, p = _preset(name)
    T = 16
    # the flight-muscle operating point: near z_line 1544 nm the connecting
    # filament sits at its rest length (params.py:1095-1099); at the skeletal
    # default of 900 nm the thick filament overruns the Z-line and a dynamic
    # lattice collapses onto the 1 nm spacing clamp
    z = 1544.0
    drv = drivers_const(1, T, z, 4.5)
    K = 5.0 * topo.n_thick
    print(f"{name} dynamic={dynamic} T={T} z={z}", flush=True)
    bt = Batch(LatticePack([topo]), np.zeros(1, np.int32), d.rows(1), drv, seeds=np.array([7], np.uint64),
               dynamic=(K, 0.0) if dynamic else None)
    bt.run(T, cfg=SolverConfig(mode="refined"))
    fg, mg = bt.force()[0], bt.metrics()[0]
    _, out = mo.run(topo, p, drv[0, :, 0], drv[0, :, 1], drv[0, :, 2], 1.0, 7,
                    mo.SolverConfig(mode="exact"), K_lat=K if dynamic else None, nu=0.0)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1)
    assert same[:4].all()
    assert out[-1, 5] > 0                      # heads bound by the end: the XB force terms are exercised
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.max(np.abs(mg[same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6
    assert np.all(mg[:, 1] <= 1e-5)


def test_collapsed_lattice_holds_the_spacing_clamp(torch_cuda):
    """Lethocerus at z_line 900 nm: the thick filament overruns the Z-line,
    titin pulls the lattice in and the spacing stops at the 1 nm clamp
    (solver.py:654). The radial residual then cannot vanish; the engine must
    report the same residual and force as the oracle, not diverge."""
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    topo, d, p = _preset("lethocerus")
    T = 4
    drv = drivers_const(1, T, 900.0, 4.5)
    K = 5.0 * topo.n_thick
    bt = Batch(LatticePack([topo]), np.zeros(1, np.int32), d.rows(1), drv, seeds=np.array([7], np.uint64),
               dynamic=(K, 0.0))
    bt.run(T, cfg=SolverConfig(mode="refined"))
    fg, mg = bt.force()[0], bt.metrics()[0]
    _, out = mo.run(topo, p, drv[0, :, 0], drv[0, :, 1], drv[0, :, 2], 1.0, 7,
                    mo.SolverConfig(mode="exact"), K_lat=K, nu=0.0)
    assert np.all(out[:, 4] == 1.0) and np.all(mg[:, 4] == 1.0)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1)
    assert same.all()
    assert np.max(np.abs(fg - out[:, 0]) / np.abs(out[:, 0])) < 1e-5
    assert np.max(np.abs(mg[:, 1] - out[:, 1]) / out[:, 1]) < 1e-4
