"""run() end to end on the GPU: the sweep it builds runs through the engine
and reshapes into the reference's data cube (simulation.py:1098-1165)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


def test_run_matches_direct_batch(torch_cuda, topo22, skeletal):
    from multifil_jax_amd import run
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    from multifil_jax_amd.simulation import plan_sweep
    s, d = skeletal
    res = run(topo22, duration_ms=12, pCa=[4.5, 6.0], replicates=2, rng_seed=3, solver="refined",
              static_params=s)
    assert res.axial_force.shape == (2, 2, 12)
    assert res._axis_names == ['pCa', 'replicates', 'time']
    assert len(res.metrics) == 52
    plan = plan_sweep(topo22, duration_ms=12, pCa=[4.5, 6.0], replicates=2)
    seeds = (np.uint64(3) << np.uint64(32)) | np.arange(4, dtype=np.uint64)
    bt = Batch(LatticePack([topo22]), np.zeros(4, np.int32), plan.params, plan.drivers, seeds=seeds)
    bt.run(12, cfg=SolverConfig(mode="refined"))
    assert np.array_equal(res.axial_force.reshape(4, 12), bt.force())
    assert np.array_equal(res.metrics['n_bound'].reshape(4, 12), bt.metrics()[:, :, 5])
    # activation: the pCa 4.5 rows bind, the pCa 6.0 rows hardly do
    assert res.sel(pCa=4.5).metrics['n_bound'][:, -1].mean() > res.sel(pCa=6.0).metrics['n_bound'][:, -1].mean()
    assert res.mean().axial_force.shape == (2, 12)


def test_run_dynamic_spacing(torch_cuda, topo22, skeletal):
    from multifil_jax_amd import run
    s, _ = skeletal
    res = run(topo22, duration_ms=8, pCa=4.5, z_line=1000.0, K_lat=5.0, solver="refined", static_params=s)
    ls = res.metrics['lattice_spacing'][0]
    assert res.topology_config['K_lat_eff'] == 5.0 * topo22.n_thick
    assert np.all(np.isfinite(ls)) and np.ptp(ls) > 0.0
    assert np.all(res.metrics['solver_residual'] <= 1e-5)


def test_autotune_picks_a_converging_config(torch_cuda, topo22, skeletal):
    from multifil_jax_amd.autotune import tune, feasible_blocks, N_CG, CG_FORCING
    from multifil_jax_amd.engine import LatticePack, Batch
    _, d = skeletal
    B, T = 32, 40
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = 1000.0, np.linspace(4.2, 6.5, B)[:, None], 14.0
    bt = Batch(LatticePack([topo22]), np.zeros(B, np.int32), d.rows(B), drv, seeds=np.arange(B, dtype=np.uint64))
    blocks = feasible_blocks(bt, "refined")
    assert 768 in blocks and 1024 in blocks
    blk, cfg = tune(bt, mode="refined", trial_steps=10, warmup_steps=5)
    assert blk in blocks and cfg.n_cg in N_CG and cfg.cg_forcing in CG_FORCING
    assert not bt.initialized.any()           # the tuner ran on a copy
    bt.block = blk
    bt.run(T, cfg=cfg)
    assert np.all(bt.metrics()[:, :, 1] <= 1e-5)


def test_cg_forcing_keeps_the_refined_tolerance(torch_cuda, topo22, skeletal):
    """The inexact-Newton CG stop changes how the solve gets there, not where.
    Free runs diverge (a 1e-8 nm difference flips a float32 nearest-site
    comparison now and then), so the check is teacher-forced: the same
    state, one step, fixed-count CG against the forcing term. The kinetics
    agree exactly; both solves meet 1e-5 pN, and their M-line forces agree
    to far below the force scale; over a free run every step meets the
    tolerance."""
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    _, d = skeletal
    B, T = 16, 30
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = np.linspace(850, 1150, B)[:, None], np.linspace(4.0, 7.0, B)[:, None], 14.0
    pack = LatticePack([topo22])
    mk = lambda: Batch(pack, np.zeros(B, np.int32), d.rows(B), drv, seeds=np.arange(B, dtype=np.uint64))
    free = mk()
    free.run(T, cfg=SolverConfig(mode="refined", cg_forcing=0.003))
    assert np.all(free.metrics()[:, :, 1] <= 1e-5)
    states = [free.get_state(b) for b in range(B)]
    out = []
    for forcing in (0.0, 0.003):
        bt = mk()
        for b in range(B):
            bt.set_state(b, states[b])
        bt.run(1, t0=T - 1, cfg=SolverConfig(mode="refined", cg_forcing=forcing))
        out.append((bt.force()[:, T - 1], bt.metrics()[:, T - 1]))
    (f0, m0), (f1, m1) = out
    assert np.array_equal(m0[:, 5:12], m1[:, 5:12]) and np.array_equal(m0[:, 19:23], m1[:, 19:23])
    assert np.all(m0[:, 1] <= 1e-5) and np.all(m1[:, 1] <= 1e-5)
    assert np.max(np.abs(f0 - f1)) < 1e-3
