"""Host side of run() / SimulationResult (simulation.py:96-420, :754-1165):
the sweep grid, driver tables, parameter rows and result reshaping, checked
without a GPU against what the reference's run() builds for the same call."""
import numpy as np
import pytest

from multifil_jax_amd.params import DYNAMIC_FIELDS, DynamicParams
from multifil_jax_amd.simulation import (METRIC_NAMES, MetricsDict, SimulationResult, plan_sweep)
from oracle import mf_oracle as mo


def test_metric_names_match_oracle():
    assert METRIC_NAMES == mo.METRIC_NAMES
    assert len(METRIC_NAMES) == 52


def test_scalar_run_is_one_instance(topo22):
    p = plan_sweep(topo22, duration_ms=10, dt=1.0)
    assert p.batch == 1 and p.n_steps == 10
    assert p.grid_shape == () and p.axis_names == ['replicates', 'time']
    assert np.all(p.drivers[0, :, 0] == 900.0) and np.all(p.drivers[0, :, 1] == 4.5)
    assert np.all(p.drivers[0, :, 2] == 14.0)
    assert np.array_equal(p.params[0], DynamicParams().row())
    assert not p.dynamic


def test_cartesian_product_and_replicates(topo22):
    # meshgrid(indexing='ij') then repeat per replicate (simulation.py:949-955)
    p = plan_sweep(topo22, duration_ms=5, pCa=[4.5, 5.0, 6.0], z_line=[900.0, 950.0], replicates=2)
    # waveform axes come in the reference's order: z_line, pCa, lattice_spacing
    assert p.grid_shape == (2, 3)
    assert p.axis_names == ['z_line', 'pCa', 'replicates', 'time']
    assert p.batch == 12
    z = p.drivers[:, 0, 0]
    pc = p.drivers[:, 0, 1]
    # z_line is the first axis (2 values), pCa the second (3 values)
    want_z = np.repeat(np.repeat([900.0, 950.0], 3), 2)
    want_pc = np.repeat(np.tile([4.5, 5.0, 6.0], 2), 2)
    assert np.allclose(z, want_z) and np.allclose(pc, want_pc)
    assert p.coords['replicates'] == [0, 1]
    assert len(p.coords['time']) == 5


def test_trace_and_trace_sweep(topo22):
    T = 8
    tr = np.linspace(900, 910, T)
    p = plan_sweep(topo22, duration_ms=T, z_line=tr)
    assert p.batch == 1 and np.allclose(p.drivers[0, :, 0], tr)
    p2 = plan_sweep(topo22, duration_ms=T, z_line=[tr, tr + 5])
    assert p2.grid_shape == (2,) and np.allclose(p2.drivers[1, :, 0], tr + 5)
    with pytest.raises(ValueError):
        plan_sweep(topo22, duration_ms=T, z_line=np.zeros(T + 1))
    with pytest.raises(ValueError):
        plan_sweep(topo22, duration_ms=T, pCa=[])


def test_dynamic_params_forms(topo22):
    p = plan_sweep(topo22, duration_ms=2, dynamic_params={'thick_k': [7000.0, 8000.0], 'thin_k': 5000.0})
    i, j = DYNAMIC_FIELDS.index('thick_k'), DYNAMIC_FIELDS.index('thin_k')
    assert p.grid_shape == (2,) and p.axis_names[0] == 'thick_k'
    assert np.allclose(p.params[:, i], [7000.0, 8000.0]) and np.all(p.params[:, j] == 5000.0)
    base = DynamicParams()
    cands = [base.copy(thick_k=7100.0), base.copy(thick_k=7200.0), base.copy(thick_k=7300.0)]
    p = plan_sweep(topo22, duration_ms=2, dynamic_params=cands)
    assert p.axis_names[0] == 'candidates' and np.allclose(p.params[:, i], [7100, 7200, 7300])
    with pytest.raises(ValueError):
        plan_sweep(topo22, duration_ms=2, dynamic_params={'no_such_field': 1.0})
    with pytest.raises(ValueError):
        plan_sweep(topo22, duration_ms=2, dynamic_params=[])
    with pytest.raises(ValueError):
        plan_sweep(topo22, duration_ms=2, dynamic_params=3.0)


def test_lattice_stiffness_and_poisson(topo22):
    # K_lat is per thick filament and scaled by n_thick (simulation.py:1002-1003)
    p = plan_sweep(topo22, duration_ms=3, K_lat=5.0)
    assert p.dynamic and np.allclose(p.K_lat, 5.0 * topo22.n_thick)
    p = plan_sweep(topo22, duration_ms=3, K_lat=[4.0, 6.0], nu=[0.0, 0.5])
    assert p.grid_shape == (2, 2) and p.axis_names[:2] == ['K_lat', 'nu']
    assert np.allclose(p.K_lat, np.repeat([4.0, 6.0], 2) * topo22.n_thick)
    # fixed spacing with nu > 0: d = d0 (z0 / z)^nu precomputed (simulation.py:1005-1012)
    z = np.array([1000.0, 1010.0, 990.0])
    p = plan_sweep(topo22, duration_ms=3, z_line=z, nu=0.5)
    assert not p.dynamic
    assert np.allclose(p.drivers[0, :, 2], 14.0 * (1000.0 / z) ** 0.5, rtol=1e-6)


def _fake_result(shape=(2, 3, 4)):
    rng = np.random.default_rng(0)
    m = MetricsDict({k: rng.normal(size=shape).astype(np.float32) for k in METRIC_NAMES})
    z = np.full(shape, 900.0)
    return SimulationResult(m, 0, z, z * 0 + 4.5, z * 0 + 14.0, dt=0.5, name="t",
                            grid_shape=shape[:-1], axis_names=['pCa', 'replicates', 'time'],
                            coords={'pCa': [4.5, 5.0], 'replicates': [0, 1, 2],
                                    'time': [0.0, 0.5, 1.0, 1.5]})


def test_result_reductions_and_slicing():
    r = _fake_result()
    assert r.n_steps == 4 and np.allclose(r.time, [0, 0.5, 1.0, 1.5])
    assert r.replicate_axis == 1
    mean = r.mean()
    assert mean.axial_force.shape == (2, 4) and 'replicates' not in mean._axis_names
    assert np.allclose(mean.axial_force, r.axial_force.mean(axis=-2))
    assert np.allclose(r.std().metrics['n_bound'], r.metrics['n_bound'].std(axis=-2))
    s = r.sel(pCa=5.0)
    assert s.axial_force.shape == (3, 4) and 'pCa' not in s.coords
    assert np.allclose(s.axial_force, r.axial_force[1])
    sl = r[:, 1:]
    assert sl.coords['replicates'] == [1, 2]
    with pytest.raises(ValueError):
        r.sel(pCa=9.0)
    with pytest.raises(ValueError):
        r.sel(nope=1)
    st = SimulationResult.stack([r, r], axis_name='lattice', axis_values=[2, 4])
    assert st.axial_force.shape == (2, 2, 3, 4) and st._axis_names[0] == 'lattice'
    assert st.name == r.name + "_stacked"                # simulation.py:410
    assert r.mean().name == r.name + "_mean" and r.std().name == r.name + "_std"
    assert abs(r.steady_state_force - float(np.mean(r.axial_force[..., -1:]))) < 1e-6
    assert "Mean force" in r.summary()
    assert r.metrics.axial_force is r.metrics['axial_force']
