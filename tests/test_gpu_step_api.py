"""The single-step API on the GPU against the FP64 oracle, teacher-forced
(same state, same injected uniforms): timestep(), kinetics_step() and
solve_equilibrium() (timestep.py:65-199, kernels/solver.py:663-771),
fixed and dynamic spacing, refined solver.

Bars: discrete state (TM, XB, bindings, nearest sites) identical; positions
within 1e-6 nm and the solved spacing within 1e-6 nm of the oracle's exact
equilibrium; residual <= 1e-5 pN; the kinetics half leaves positions
unchanged bit for bit; successive keys differ.
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo
from oracle.philox import uniforms, PURPOSE_TM, PURPOSE_XB

from gpu_helpers import oracle_trajectory

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    return torch


def _api_state(st, topo, ls=14.0):
    from multifil_jax_amd.timestep import State, ThickState, ThinState
    shp = (topo.n_thick, topo.n_crowns, topo.n_xb_per_crown)
    dist = np.stack([st.xb_xdist, np.full(st.xb_xdist.shape, ls, np.float32)], -1).reshape(shp + (2,))
    return State(ThickState(st.thick.copy(), st.xb_state.reshape(shp).copy(), st.xb_bound.reshape(shp).copy(),
                            st.xb_nearest.reshape(shp).copy(), dist),
                 ThinState(st.thin.copy(), st.tm.copy(), st.thin_bound.copy()))


def _check_discrete(got, want):
    assert np.array_equal(got.thick.xb_states.ravel(), want.xb_state)
    assert np.array_equal(got.thick.xb_bound_to.ravel(), want.xb_bound)
    assert np.array_equal(got.thin.bound_to, want.thin_bound)
    assert np.array_equal(got.thin.tm_states, want.tm)
    assert np.array_equal(got.thick.xb_nearest_bs.ravel(), want.xb_nearest)
    assert np.array_equal(got.thick.xb_distances[..., 0].ravel(), want.xb_xdist)


def test_timestep_fixed_spacing(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd.timestep import timestep, Drivers
    _, d = skeletal
    states, tab = oracle_trajectory(topo22, pdict, 1000.0, 4.5, 8, seed=11)
    S, X = topo22.n_thin * topo22.n_sites, topo22.total_xbs
    pre = mo.preconditioner(topo22, pdict)
    key = np.uint64(1234)
    for k in (4, 8):
        st = states[k]
        u_tm, u_xb = uniforms(500 + k, 0, PURPOSE_TM, S), uniforms(500 + k, 0, PURPOSE_XB, X)
        want, _ = mo.step(st, topo22, pdict, tab, 1000.0, 4.5, 14.0, 0.0, u_tm, u_xb,
                          mo.SolverConfig(mode="exact"), pre)
        print(f"timestep from oracle state {k}", flush=True)
        got, key2, res, ls, it = timestep(_api_state(st, topo22), d, Drivers(4.5, 1000.0, np.nan), topo22, key,
                                          dt=1.0, uniforms=(u_tm, u_xb), solver="refined")
        _check_discrete(got, want)
        assert np.max(np.abs(got.thick.axial.ravel() - want.thick.ravel())) < 1e-6
        assert np.max(np.abs(got.thin.axial.ravel() - want.thin.ravel())) < 1e-6
        assert res <= 1e-5 and ls == 14.0 and it >= 1
        assert key2 != key
        key = key2


def test_kinetics_then_solve(torch_cuda, topo22, skeletal, pdict):
    """The split halves compose to the oracle's step: kinetics_step moves no
    node, solve_equilibrium reaches the oracle's equilibrium."""
    from multifil_jax_amd.timestep import kinetics_step, solve_equilibrium, Drivers
    _, d = skeletal
    states, tab = oracle_trajectory(topo22, pdict, 1000.0, 4.5, 6, seed=12)
    S, X = topo22.n_thin * topo22.n_sites, topo22.total_xbs
    st = states[6]
    u_tm, u_xb = uniforms(77, 0, PURPOSE_TM, S), uniforms(77, 0, PURPOSE_XB, X)
    mid = mo.kinetics(st, topo22, tab, u_tm, u_xb)
    api = _api_state(st, topo22)
    print("kinetics_step", flush=True)
    k_state, _, resolved = kinetics_step(api, d, Drivers(np.nan, 1000.0, np.nan), topo22, np.uint64(5),
                                         dt=1.0, uniforms=(u_tm, u_xb), solver="refined")
    _check_discrete(k_state, mid)
    assert np.array_equal(k_state.thick.axial, api.thick.axial)
    assert np.array_equal(k_state.thin.axial, api.thin.axial)
    assert float(resolved.pCa) == np.float32(d.pCa) and float(resolved.z_line) == 1000.0
    print("solve_equilibrium", flush=True)
    x, _, _, _ = mo.solve(mid, topo22, pdict, 1000.0, 14.0, mo.SolverConfig(mode="exact"),
                          mo.preconditioner(topo22, pdict))
    new, res, ls, it = solve_equilibrium(k_state, resolved, topo22, solver="refined")
    print(f"  residual {res:.2e}, {it} Newton iterations", flush=True)
    assert (mid.xb_bound >= 0).any()          # the solve has bound heads to balance
    assert np.max(np.abs(np.concatenate([new.thick.axial.ravel(), new.thin.axial.ravel()]) - x)) < 1e-6
    assert res <= 1e-5 and ls == 14.0 and it >= 1
    assert np.array_equal(new.thick.xb_nearest_bs, k_state.thick.xb_nearest_bs)
    # reference solver: the reference's caps and floor
    new_r, res_r, _, it_r = solve_equilibrium(k_state, resolved, topo22, solver="reference", n_newton_steps=4)
    assert res_r <= mo.solver_tolerance(pdict) + 0.6 or it_r == 4


def test_timestep_dynamic_spacing(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd.timestep import timestep, Drivers
    _, d = skeletal
    states, _ = oracle_trajectory(topo22, pdict, 1000.0, 4.5, 6, seed=13)
    S, X = topo22.n_thin * topo22.n_sites, topo22.total_xbs
    K = 5.0 * topo22.n_thick
    st = states[6].copy()
    ls0 = 14.2                      # the carried spacing: kinetics spacing and first guess
    st.ls = ls0
    tab = mo.make_tables(pdict, 4.5, ls0, 1.0, topo22.xb_bin_centers)
    u_tm, u_xb = uniforms(88, 0, PURPOSE_TM, S), uniforms(88, 0, PURPOSE_XB, X)
    want, mt = mo.step(st, topo22, pdict, tab, 1000.0, 4.5, ls0, 0.0, u_tm, u_xb, mo.SolverConfig(mode="exact"),
                       mo.preconditioner(topo22, pdict), dynamic=(K, 14.0))
    print("timestep, dynamic spacing", flush=True)
    got, _, res, ls, it = timestep(_api_state(st, topo22, ls0), d, Drivers(4.5, 1000.0, ls0), topo22,
                                   np.uint64(9), dt=1.0, K_lat=K, d_ref=14.0, uniforms=(u_tm, u_xb),
                                   solver="refined")
    _check_discrete(got, want)
    assert abs(ls - want.ls) < 1e-6
    assert np.max(np.abs(got.thick.axial.ravel() - want.thick.ravel())) < 1e-6
    assert res <= 1e-5


def test_create_and_equilibrate(torch_cuda, topo22, skeletal, pdict):
    from multifil_jax_amd.timestep import create_and_equilibrate
    _, d = skeletal
    got = create_and_equilibrate(topo22, d, 1000.0, solver="refined")
    want = mo.init_state(topo22, pdict, 1000.0, mo.SolverConfig(mode="exact"))
    x = np.concatenate([got.thick.axial.ravel(), got.thin.axial.ravel()])
    assert np.max(np.abs(x - want.x())) < 1e-6
    assert np.array_equal(got.thick.xb_nearest_bs.ravel(), want.xb_nearest)
