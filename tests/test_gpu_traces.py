"""Time-varying drivers on the GPU against the FP64 oracle, through run().

* A calcium twitch: pCa(t) = 9 - 5 exp(-((t - 50) / 20)^2 / 2) at z_line
  1100 nm (examples/quickstart.py:62-64), 80 steps of dt 1 ms, so the
  transition tables change at every step.
* A fixed-spacing Poisson trace: a Z-line ramp with nu = 0.5 and no K_lat,
  so run() turns the scalar spacing into d0 (z0 / z)^nu per step
  (simulation.py:1005-1012) and the XB tables change at every step.
* NaN driver entries fall back to the parameter row's value
  (core/state.py:331-343): a trace with NaN holes runs exactly like the
  trace with the holes filled by the constant.

Bars: refined solver; while the discrete state counts agree (required for
at least the first 30 steps), the M-line force is within 1e-5 relative of
the FP64 oracle fed the same Philox draws, and the pCa / spacing metric
columns equal the drivers.
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


def _twitch(T):
    t = np.arange(T, dtype=np.float32)
    return (9.0 - 5.0 * np.exp(-0.5 * ((t - 50.0) / 20.0) ** 2)).astype(np.float32)


def _compare(res, plan, topo, pdict, seed, min_same=30):
    drv = plan.drivers[0]
    _, out = mo.run(topo, pdict, drv[:, 0], drv[:, 1], drv[:, 2], 1.0, seed, mo.SolverConfig(mode="exact"))
    mg = np.stack([res.metrics[k].reshape(-1) for k in mo.METRIC_NAMES], axis=1)
    fg = res.axial_force.reshape(-1)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1) & np.all(mg[:, 19:23] == out[:, 19:23], axis=1)
    n_same = int(np.argmin(same)) if not same.all() else len(same)
    print(f"  steps with identical state counts: {n_same} of {len(same)}", flush=True)
    assert n_same >= min_same
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.all(mg[:, 1] <= 1e-5)
    return mg, out


def test_pca_twitch_matches_oracle(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd import run
    from multifil_jax_amd.simulation import plan_sweep
    s, d = skeletal
    T = 80
    pca = _twitch(T)
    print("pCa twitch, 2x2, 80 steps", flush=True)
    res = run(topo22, pCa=pca, z_line=1100.0, duration_ms=T, dt=1.0, rng_seed=0, solver="refined",
              static_params=s, dynamic_params=d)
    plan = plan_sweep(topo22, duration_ms=T, pCa=pca, z_line=1100.0, dynamic_params=d)
    mg, out = _compare(res, plan, topo22, pdict, seed=0)
    assert np.array_equal(mg[:, 3], pca)
    # the twitch activates and binds heads around its calcium peak
    assert out[:, 5].max() > 0 and mg[:, 5].max() > 0
    assert mg[:20, 5].max() == 0


def test_poisson_spacing_trace_matches_oracle(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd import run
    from multifil_jax_amd.simulation import plan_sweep
    s, d = skeletal
    T = 40
    z = np.linspace(1000.0, 1060.0, T, dtype=np.float32)
    print("fixed-spacing Poisson trace, 2x2, 40 steps", flush=True)
    res = run(topo22, pCa=4.5, z_line=z, lattice_spacing=14.0, nu=0.5, duration_ms=T, dt=1.0,
              rng_seed=1, solver="refined", static_params=s, dynamic_params=d)
    plan = plan_sweep(topo22, duration_ms=T, pCa=4.5, z_line=z, lattice_spacing=14.0, nu=0.5,
                      dynamic_params=d)
    ls = plan.drivers[0, :, 2]
    want = np.float32(14.0) * (z[0] / z) ** np.float32(0.5)
    assert np.allclose(ls, want, rtol=1e-6) and len(np.unique(ls)) == T
    mg, _ = _compare(res, plan, topo22, pdict, seed=(1 << 32) | 0)
    assert np.array_equal(mg[:, 4], ls)


def test_nan_driver_sentinel(torch_cuda, topo22, skeletal):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    _, d = skeletal
    T = 12
    drv = np.zeros((2, T, 3), np.float32)
    drv[:, :, 0] = 1000.0
    drv[:, :, 1] = np.linspace(5.0, 4.5, T, dtype=np.float32)
    drv[:, :, 2] = 14.0
    holes = drv.copy()
    holes[0, 3:6, 1] = np.nan           # pCa holes -> the row's pCa (4.5)
    holes[1, 4:9, 2] = np.nan           # spacing holes -> the row's spacing (14)
    filled = drv.copy()
    filled[0, 3:6, 1] = np.float32(d.pCa)
    filled[1, 4:9, 2] = np.float32(d.lattice_spacing)
    cfg = SolverConfig(mode="refined")
    runs = []
    for trace in (holes, filled):
        bt = Batch(LatticePack([topo22]), np.zeros(2, np.int32), d.rows(2), trace,
                   seeds=np.array([4, 5], np.uint64))
        bt.run(T, cfg=cfg)
        runs.append((bt.force(), bt.metrics()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert not np.isnan(runs[0][1]).any()


@pytest.mark.parametrize("dynamic", [False, True])
def test_in_kernel_rebuild_equals_split_launches(torch_cuda, topo22, skeletal, dynamic):
    """A pCa twitch and a spacing ramp in one launch, each instance
    rebuilding its own tables in the kernel when its drivers change
    (MF_LAUNCH_REBUILD), give the same bits as one launch per constant
    stretch with mf_tables between them (three instances: a twitch, a
    spacing ramp, a constant one).  Both runs use the dynamic-spacing
    kernel variants, which carry the in-kernel rebuild, so they differ only
    in where the tables come from (another kernel variant may round the
    solve differently in the last FP64 bits).  With dynamic spacing the
    kernel rebuilds at every step either way; there the check covers the
    rebuild of only the rows a step reads."""
    from multifil_jax_amd import _lib
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    _, d = skeletal
    T = 60
    drv = np.zeros((3, T, 3), np.float32)
    drv[:, :, 0] = 1000.0
    drv[:, :, 2] = 14.0
    drv[0, :, 1] = _twitch(T + 20)[20:]
    drv[1, :, 1] = 4.8
    drv[1, :, 2] = np.linspace(13.5, 14.5, T, dtype=np.float32)
    drv[2, :, 1] = 5.2
    K = 5.0 * topo22.n_thick
    out = []
    for split in (False, True):
        print(f"twitch / spacing ramp, split={split}, dynamic={dynamic}", flush=True)
        bt = Batch(LatticePack([topo22]), np.zeros(3, np.int32), d.rows(3), drv, seeds=np.arange(3, dtype=np.uint64),
                   dynamic=(np.full(3, K, np.float32), np.zeros(3, np.float32)) if dynamic else None)
        bt.run(T, cfg=SolverConfig(mode="refined"), split=split,
               launch_flags=_lib.MF_LAUNCH_DYNAMIC if split else 0)
        out.append((bt.force(), bt.metrics(), [bt.get_state(b) for b in range(3)]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1], equal_nan=True)
    for a, b in zip(out[0][2], out[1][2]):
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert out[0][1][0, :, 5].max() > 0          # the twitch binds heads


@pytest.mark.parametrize("mode", ["mean_field", "c_zone"])
def test_in_kernel_rebuild_with_subpopulations_equals_split_launches(torch_cuda, topo22, skeletal, mode):
    """As above with subpopulations on a pCa twitch (fixed spacing): a
    mean-field fraction, whose blended TM and XB rows, or c_zone labels,
    whose per-label row sets, each instance rebuilds in the kernel when its
    pCa changes; the same bits as one launch per constant stretch."""
    from multifil_jax_amd import _lib, Subpopulation
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig, Populations
    from multifil_jax_amd.simulation import plan_sweep, _populations
    _, d = skeletal
    sc = dict(xb_r01_coeff=3.0, tm_k_01=0.5)
    sp = Subpopulation.c_zone(topo22, **sc) if mode == "c_zone" else Subpopulation.mean_field(0.4, **sc)
    T = 60
    plan = plan_sweep(topo22, duration_ms=T, pCa=[4.8, 5.2], z_line=1000.0, dynamic_params=d, subpopulation=sp)
    pops = _populations(plan, Populations)
    drv = np.array(plan.drivers, np.float32)
    drv[0, :, 1] = _twitch(T + 20)[20:]
    out = []
    for split in (False, True):
        print(f"twitch with {mode} subpopulations, split={split}", flush=True)
        bt = Batch(LatticePack([topo22]), np.zeros(2, np.int32), plan.params, drv,
                   seeds=np.arange(2, dtype=np.uint64), pops=pops)
        bt.run(T, cfg=SolverConfig(mode="refined"), split=split,
               launch_flags=_lib.MF_LAUNCH_DYNAMIC if split else 0)
        out.append((bt.force(), bt.metrics()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1], equal_nan=True)
    assert out[0][1][0, :, 5].max() > 0
