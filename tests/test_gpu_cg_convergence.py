"""The solve's CG converges as fast as the exact backbone preconditioner
allows.

A preconditioner that is wrong but still symmetric positive definite only
slows CG: the FP64 Newton residual still drives the solve to the right
equilibrium, so no parity test sees it.  This test does.  From activated
post-kinetics oracle states (2x2, 3x3, 4x4 lattices; one variant each of
the 2, 5 and 8 node slots per lane), one refined solve_equilibrium on the
GPU runs with at most 2 CG passes per Newton iteration and no forcing term,
and the oracle runs the same schedule on the CPU: Newton with 2 unrolled
PCG iterations per step on its Thomas-factored backbone preconditioner
(solver.py:243-366), stopping at max|F| <= 1e-5 pN.  In exact arithmetic
the two produce the same iterates (the device's Chronopoulos-Gear CG
equals textbook PCG), so the device needs no more Newton iterations than
the oracle.  A preconditioner off by the padding-slot sums of an 86-site
thin chain doubles the count (2x2: 3 -> 6, 3x3: 2 -> 5, 4x4: 3 -> 5).
"""
import numpy as np
import pytest

from oracle import mf_oracle as mo
from oracle.philox import uniforms, PURPOSE_TM, PURPOSE_XB

from gpu_helpers import oracle_trajectory

pytestmark = pytest.mark.gpu

K_CG = 2


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


@pytest.mark.parametrize("n,pCa,steps", [(2, 4.4, 25), (3, 4.2, 12), (4, 4.2, 12)])
def test_cg_rate_matches_exact_preconditioner(torch_cuda, skeletal, pdict, n, pCa, steps):
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.timestep import solve_equilibrium
    s, d = skeletal
    topo = SarcTopology.create(n, n, s, d)
    z = pdict["z_line"]
    states, tab = oracle_trajectory(topo, pdict, z, pCa, steps, seed=3)
    S, X = topo.n_thin * topo.n_sites, topo.total_xbs
    pre = mo.preconditioner(topo, pdict)
    for k in (steps - 1, steps):
        st = mo.kinetics(states[k], topo, tab, uniforms(90 + k, 0, PURPOSE_TM, S),
                         uniforms(90 + k, 0, PURPOSE_XB, X))
        m = mo.Mech(topo, pdict, z, 14.0, st.xb_state, st.xb_bound)
        x0 = st.x()
        f0 = np.abs(mo.residual(m, x0)).max()
        _, it_o, res_o = mo.newton_cg(m, x0, pre, 1e-5, 32, K_CG)
        print(f"{n}x{n} state {k}: bound {int(m.bound.sum())}, |F0| {f0:.2f} pN, "
              f"oracle {it_o} Newton iterations", flush=True)
        new, res, ls, it = solve_equilibrium(_api_state(st, topo), d, topo, solver="refined",
                                             n_cg_steps=K_CG, n_newton_steps=16)
        print(f"  device {it} Newton iterations, residual {res:.2e}", flush=True)
        assert f0 > 1.0 and res_o <= 1e-5
        assert res <= 1e-5
        assert it <= it_o
