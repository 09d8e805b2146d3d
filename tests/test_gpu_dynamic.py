"""GPU parity of dynamic lattice spacing and uneven batches against the FP64
oracle (bordered Newton system of solver.py:599-656, K_lat / nu drivers of
simulation.py:691-696, metrics at the solved spacing of :711-719).

Bars (stated per check):
* refined solver, free run with identical Philox draws: while the XB state
  counts agree, the M-line force is within 1e-5 relative and the solved
  spacing within 1e-6 nm of the FP64 oracle's exact bordered solve; the
  expected-ATP / tear metrics (absorbing chain at the solved spacing) within
  2e-5 relative + 1e-5;
* reference solver: residual (axial and radial) within the reference's
  floor or the iteration cap; spacing within 0.02 nm of the exact solve
  (the 0.75 pN stopping floor over a radial stiffness of order 100 pN/nm
  allows ~0.01 nm);
* a fixed-spacing instance run by the dynamic-spacing kernel matches the
  fixed-spacing kernel: identical discrete state, force within 1e-9
  relative (two compiled kernels may contract multiply-adds differently);
* uneven lengths and mixed lattices in one launch reproduce per-instance
  separate runs at the same block size (same discrete state, force within
  1e-9 relative), and rows past an instance's length stay zero.
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo

from gpu_helpers import drivers_const

pytestmark = pytest.mark.gpu

K_LAT = 5.0   # pN/nm per thick filament (examples/benchmarks/benchmark_dynamic_ls.py:128-145)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


def _batch(topos, ids, d, drv, seeds, dynamic=None, lengths=None):
    from multifil_jax_amd.engine import LatticePack, Batch
    B = len(ids)
    return Batch(LatticePack(topos), np.asarray(ids, np.int32), d.rows(B), drv,
                 seeds=np.asarray(seeds, np.uint64), dynamic=dynamic, lengths=lengths)


def _cfg(mode):
    from multifil_jax_amd.engine import SolverConfig
    return SolverConfig(mode=mode)


@pytest.mark.parametrize("block", [None, 768, 1024])
def test_dynamic_free_run_refined(torch_cuda, topo22, skeletal, pdict, block):
    """Default block (512 threads: three node slots per lane, the in-kernel
    rebuild by lane pairs) and 768 / 1024 threads (two slots, lane quads)."""
    _, d = skeletal
    print(f"2x2 dynamic free run, block={block}", flush=True)
    T = 20
    K = K_LAT * topo22.n_thick
    # instance 0: constant Z-line, nu = 0; instance 1: a 10 nm ramp with nu = 0.5
    drv = drivers_const(2, T, np.array([1000.0, 1000.0], np.float32), np.array([4.5, 5.0], np.float32))
    drv[1, :, 0] = np.linspace(1000.0, 1010.0, T, dtype=np.float32)
    nu = np.array([0.0, 0.5], np.float32)
    bt = _batch([topo22], [0, 0], d, drv, [7, 8], dynamic=(np.full(2, K, np.float32), nu))
    bt.run(T, cfg=_cfg("refined"), block=block)
    assert bt.block == (block or 512)
    fg, mg = bt.force(), bt.metrics()
    for b in range(2):
        _, out = mo.run(topo22, pdict, drv[b, :, 0], drv[b, :, 1], drv[b, :, 2], 1.0, 7 + b,
                        mo.SolverConfig(mode="exact"), K_lat=K, nu=float(nu[b]))
        same = np.all(mg[b, :, 6:12] == out[:, 6:12], axis=1)
        assert same[:6].all(), f"instance {b}: state counts diverge early"
        assert np.max(np.abs(fg[b] - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
        assert np.max(np.abs(mg[b, same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6   # float32 metric
        assert np.all((mg[b, :, 4] > 5.0) & (mg[b, :, 4] < 25.0))
        for col in (48, 49):     # atp_expected_p, xb_tear_expected at the solved spacing
            assert np.all(np.abs(mg[b, same, col] - out[same, col])
                          <= 2e-5 * np.abs(out[same, col]) + 1e-5), col
        assert np.all(mg[b, :, 1] <= 1e-5)


def test_dynamic_reference_mode(torch_cuda, topo22, skeletal, pdict):
    _, d = skeletal
    T = 15
    K = K_LAT * topo22.n_thick
    drv = drivers_const(1, T, 1000.0, 4.5)
    bt = _batch([topo22], [0], d, drv, [3], dynamic=(K, 0.0))
    bt.run(T, cfg=_cfg("reference"))
    mg = bt.metrics()[0]
    tol = mo.solver_tolerance(pdict)
    assert np.all((mg[:, 1] <= tol + 0.6) | (mg[:, 51] == 4))
    _, out = mo.run(topo22, pdict, drv[0, :, 0], drv[0, :, 1], drv[0, :, 2], 1.0, 3,
                    mo.SolverConfig(mode="exact"), K_lat=K, nu=0.0)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1)
    assert same[:4].all()
    assert np.max(np.abs(mg[same, 4] - out[same, 4])) < 0.02


def test_fixed_instance_inside_dynamic_launch(torch_cuda, topo22, skeletal):
    _, d = skeletal
    T = 10
    drv = drivers_const(2, T, np.array([950.0, 1050.0], np.float32), np.array([4.5, 5.5], np.float32))
    K = np.array([K_LAT * topo22.n_thick, np.nan], np.float32)
    a = _batch([topo22], [0, 0], d, drv, [1, 2], dynamic=(K, np.zeros(2, np.float32)))
    a.run(T, cfg=_cfg("refined"))
    b = _batch([topo22], [0], d, drv[1:], [2])
    b.run(T, cfg=_cfg("refined"))
    assert np.allclose(a.force()[1], b.force()[0], rtol=1e-9, atol=0)
    assert np.array_equal(a.metrics()[1, :, 5:12], b.metrics()[0, :, 5:12])
    assert np.all(a.metrics()[1, :, 4] == 14.0)


def test_uneven_lengths_and_mixed_lattices(torch_cuda, skeletal):
    """2x2, 3x3 and 4x4 instances of different lengths in one dynamic launch
    give the same trajectories as separate single-lattice launches."""
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    topos = [SarcTopology.create(n, n, s, d) for n in (2, 3, 4)]
    T = 12
    ids = [0, 1, 2, 0]
    lengths = np.array([12, 7, 9, 5])
    B = len(ids)
    drv = drivers_const(B, T, np.array([1000, 1020, 980, 1000], np.float32),
                        np.array([4.6, 5.2, 4.8, 6.0], np.float32))
    K = np.array([K_LAT * topos[i].n_thick for i in ids], np.float32)
    seeds = np.arange(10, 10 + B)
    mixed = _batch(topos, ids, d, drv, seeds, dynamic=(K, np.zeros(B, np.float32)), lengths=lengths)
    mixed.run(T, cfg=_cfg("refined"))
    fm, mm = mixed.force(), mixed.metrics()
    for b in range(B):
        one = _batch([topos[ids[b]]], [0], d, drv[b:b + 1, :lengths[b]], seeds[b:b + 1],
                     dynamic=(K[b:b + 1], np.zeros(1, np.float32)))
        one.run(int(lengths[b]), cfg=_cfg("refined"), block=mixed.block)
        L = int(lengths[b])
        assert np.allclose(fm[b, :L], one.force()[0], rtol=1e-9, atol=0), b
        assert np.array_equal(mm[b, :L, 5:12], one.metrics()[0, :, 5:12]), b
        assert np.all(mm[b, L:] == 0.0) and np.all(fm[b, L:] == 0.0), b


@pytest.mark.parametrize("lat", [3, 4])
def test_dynamic_free_run_refined_3x3_4x4(torch_cuda, skeletal, pdict, lat):
    """The bars of the 2x2 free run at the uneven batch's middle lattices:
    3x3 (odd rows, five node slots per lane) and 4x4, one instance at a
    constant Z-line and one on a 10 nm ramp with nu = 0.5."""
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    topo = SarcTopology.create(lat, lat, s, d)
    T = 24
    K = K_LAT * topo.n_thick
    drv = drivers_const(2, T, np.array([1000.0, 1000.0], np.float32), np.array([4.5, 5.2], np.float32))
    drv[1, :, 0] = np.linspace(1000.0, 1010.0, T, dtype=np.float32)
    nu = np.array([0.0, 0.5], np.float32)
    print(f"dynamic free run {lat}x{lat}", flush=True)
    bt = _batch([topo], [0, 0], d, drv, [21, 22], dynamic=(np.full(2, K, np.float32), nu))
    bt.run(T, cfg=_cfg("refined"))
    fg, mg = bt.force(), bt.metrics()
    for b in range(2):
        _, out = mo.run(topo, pdict, drv[b, :, 0], drv[b, :, 1], drv[b, :, 2], 1.0, 21 + b,
                        mo.SolverConfig(mode="exact"), K_lat=K, nu=float(nu[b]))
        same = np.all(mg[b, :, 6:12] == out[:, 6:12], axis=1)
        assert same[:6].all(), f"instance {b}: state counts diverge early"
        assert np.max(np.abs(fg[b] - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
        assert np.max(np.abs(mg[b, same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6
        for col in (48, 49):
            assert np.all(np.abs(mg[b, same, col] - out[same, col])
                          <= 2e-5 * np.abs(out[same, col]) + 1e-5), col
        assert np.all(mg[b, :, 1] <= 1e-5)
        assert out[-1, 5] > 0
