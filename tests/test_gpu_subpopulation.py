"""Subpopulations on the GPU (core/subpopulation.py; transitions.py:580-602,
925-975), through run().

Bars:
* the reference's exact edges (subpopulation.py:51-56), bit for bit: a
  mean-field fraction of 0 reproduces the wild-type run, a fraction of 1
  the run with the scales applied globally; explicit labels that are all 1
  reproduce the global scale too;
* parity against the FP64 oracle given the same labels and Philox draws, in
  all three modes, refined solver: the discrete state counts agree, and
  while they agree the M-line force is within 1e-5 relative.
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo

pytestmark = pytest.mark.gpu

SCALES = dict(xb_r01_coeff=3.0, tm_k_01=0.5)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


def _run(topo, s, **kw):
    from multifil_jax_amd import run
    return run(topo, duration_ms=10, pCa=[4.5, 5.5], z_line=950.0, rng_seed=3, solver="refined",
               static_params=s, **kw)


def _same(a, b):
    assert np.array_equal(a.axial_force, b.axial_force)
    for k in a.metrics:
        assert np.array_equal(a.metrics[k], b.metrics[k], equal_nan=True), k


def _scaled(d):
    return d.copy(**{k: np.float32(getattr(d, k)) * np.float32(v) for k, v in SCALES.items()})


def test_mean_field_edges_are_bit_exact(torch_cuda, topo22, skeletal):
    from multifil_jax_amd import Subpopulation
    s, d = skeletal
    base = _run(topo22, s, dynamic_params=d)
    print("mean-field fraction 0", flush=True)
    _same(base, _run(topo22, s, dynamic_params=d, subpopulation=Subpopulation.mean_field(0.0, **SCALES)))
    glob = _run(topo22, s, dynamic_params=_scaled(d))
    print("mean-field fraction 1", flush=True)
    _same(glob, _run(topo22, s, dynamic_params=d, subpopulation=Subpopulation.mean_field(1.0, **SCALES)))
    assert not np.array_equal(base.metrics['n_bound'], glob.metrics['n_bound'])


def test_explicit_all_mutant_is_the_global_scale(torch_cuda, topo22, skeletal):
    from multifil_jax_amd import Subpopulation
    s, d = skeletal
    glob = _run(topo22, s, dynamic_params=_scaled(d))
    print("random fraction 1", flush=True)
    _same(glob, _run(topo22, s, dynamic_params=d, subpopulation=Subpopulation.random(1.0, seed=4, **SCALES)))


# Scale sets: both kinetics (the XB tables do not fit LDS beside a second
# XB set at 2x2), only TM (one XB set, copied into LDS after the two TM
# sets) and only XB.
SCALE_SETS = {"both": SCALES, "tm_only": dict(tm_k_01=0.5), "xb_only": dict(xb_r01_coeff=3.0)}


@pytest.mark.parametrize("mode,scales", [("mean_field", "both"), ("random", "both"), ("c_zone", "both"),
                                         ("random", "tm_only"), ("c_zone", "tm_only"),
                                         ("random", "xb_only")])
def test_subpopulation_parity(torch_cuda, topo22, skeletal, pdict, mode, scales):
    from multifil_jax_amd import Subpopulation
    from multifil_jax_amd.simulation import plan_sweep
    from multifil_jax_amd.subpopulation import population_params
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    s, d = skeletal
    sc = SCALE_SETS[scales]
    sp = {"mean_field": lambda: Subpopulation.mean_field(0.4, **sc),
          "random": lambda: Subpopulation.random(0.5, seed=9, **sc),
          "c_zone": lambda: Subpopulation.c_zone(topo22, **sc)}[mode]()
    T = 12
    print(f"subpopulation parity {mode} {scales}", flush=True)
    from multifil_jax_amd import run
    res = run(topo22, duration_ms=T, pCa=4.5, z_line=950.0, rng_seed=3, solver="refined",
              static_params=s, dynamic_params=d, subpopulation=sp)
    plan = plan_sweep(topo22, duration_ms=T, pCa=4.5, z_line=950.0, dynamic_params=d, subpopulation=sp)
    rp = plan.subpop
    P = population_params(plan.params, rp)[0]
    pops = [{n: float(P[k, i]) for i, n in enumerate(DYNAMIC_FIELDS)} for k in range(rp.K)]
    pop = mo.Pop(rp.mode, pops, rp.fractions[0], has_xb=rp.has_xb, has_tm=rp.has_tm,
                 xb_label=None if rp.xb_mask is None else rp.xb_mask[0],
                 tm_label=None if rp.tm_mask is None else rp.tm_mask[0])
    if mode != "mean_field":
        mask = rp.xb_mask[0] if rp.has_xb else rp.tm_mask[0]
        assert 0 < mask.sum() < mask.size
    seed = (3 << 32) | 0
    drv = plan.drivers[0]
    _, out = mo.run(topo22, pdict, drv[:, 0], drv[:, 1], drv[:, 2], 1.0, seed,
                    mo.SolverConfig(mode="exact"), pop=pop)
    mg = np.stack([res.metrics[k].reshape(-1) for k in mo.METRIC_NAMES], axis=1)
    fg = res.axial_force.reshape(-1)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1) & np.all(mg[:, 19:23] == out[:, 19:23], axis=1)
    assert same[:6].all()
    assert out[-1, 5] > 0
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.all(mg[:, 1] <= 1e-5)


@pytest.mark.parametrize("mode", ["c_zone", "mean_field"])
def test_subpopulation_with_dynamic_spacing(torch_cuda, topo22, skeletal, pdict, mode):
    """Subpopulations on a dynamic-spacing run: c_zone labels (XB and TM),
    whose per-label tables, or a mean-field fraction of 0.4, whose blended
    generators, are rebuilt in the kernel at the carried and at the solved
    spacing each step; parity against the oracle as above, plus the solved
    spacing within 1e-6 nm."""
    from multifil_jax_amd import Subpopulation, run
    from multifil_jax_amd.simulation import plan_sweep
    from multifil_jax_amd.subpopulation import population_params
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    s, d = skeletal
    sp = (Subpopulation.c_zone(topo22, **SCALES) if mode == "c_zone"
          else Subpopulation.mean_field(0.4, **SCALES))
    T, K = 12, 5.0
    print(f"subpopulation {mode} with dynamic spacing", flush=True)
    res = run(topo22, duration_ms=T, pCa=4.5, z_line=950.0, rng_seed=3, solver="refined",
              static_params=s, dynamic_params=d, subpopulation=sp, K_lat=K)
    plan = plan_sweep(topo22, duration_ms=T, pCa=4.5, z_line=950.0, dynamic_params=d, subpopulation=sp,
                      K_lat=K)
    rp = plan.subpop
    P = population_params(plan.params, rp)[0]
    pops = [{n: float(P[k, i]) for i, n in enumerate(DYNAMIC_FIELDS)} for k in range(rp.K)]
    pop = mo.Pop(rp.mode, pops, rp.fractions[0], has_xb=rp.has_xb, has_tm=rp.has_tm,
                 xb_label=None if rp.xb_mask is None else rp.xb_mask[0],
                 tm_label=None if rp.tm_mask is None else rp.tm_mask[0])
    drv = plan.drivers[0]
    _, out = mo.run(topo22, pdict, drv[:, 0], drv[:, 1], drv[:, 2], 1.0, (3 << 32) | 0,
                    mo.SolverConfig(mode="exact"), K_lat=float(plan.K_lat[0]), nu=0.0, pop=pop)
    mg = np.stack([res.metrics[k].reshape(-1) for k in mo.METRIC_NAMES], axis=1)
    fg = res.axial_force.reshape(-1)
    same = np.all(mg[:, 6:12] == out[:, 6:12], axis=1) & np.all(mg[:, 19:23] == out[:, 19:23], axis=1)
    assert same[:6].all()
    assert np.max(np.abs(fg - out[:, 0])[same] / np.abs(out[same, 0])) < 1e-5
    assert np.max(np.abs(mg[same, 4] - out[same, 4])) < 1e-6 * 14 + 2e-6
    for col in (48, 49):     # expected ATP / tear from the per-label absorbing rows
        assert np.all(np.abs(mg[same, col] - out[same, col]) <= 2e-5 * np.abs(out[same, col]) + 1e-5), col
    assert np.all(mg[:, 1] <= 1e-5)
