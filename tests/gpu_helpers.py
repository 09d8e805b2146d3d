"""Shared helpers for the GPU parity tests (test infrastructure only)."""
import numpy as np

from oracle import mf_oracle as mo
from oracle.philox import uniforms, PURPOSE_TM, PURPOSE_XB

# Per-instance table layout written by mf_tables (floats).
TAB_TM = 16
TAB_XB = 16 + 27 * 16


def tab_abs(nb):
    return TAB_XB + 2 * nb * 48


def oracle_state_to_device(st: mo.State, topo) -> dict:
    key = mo.xb_keys(st.xb_xdist, st.xb_nearest, st.tm, topo)
    return dict(pos=st.x(), ls=st.ls, xb_bound=st.xb_bound.astype(np.int32),
                xb_key=key.astype(np.int32), thin_bound=st.thin_bound.ravel().astype(np.int32),
                xb_state=st.xb_state.astype(np.uint8), tm=st.tm.ravel().astype(np.uint8))


def make_batch(topos, lattice_ids, params_rows, drivers, seeds):
    from multifil_jax_amd.engine import LatticePack, Batch
    return Batch(LatticePack(topos), np.asarray(lattice_ids), params_rows, drivers, seeds)


def drivers_const(B, T, z, pCa, ls=14.0):
    d = np.zeros((B, T, 3), np.float32)
    d[:, :, 0] = np.asarray(z, np.float32).reshape(-1, 1) if np.ndim(z) else z
    d[:, :, 1] = np.asarray(pCa, np.float32).reshape(-1, 1) if np.ndim(pCa) else pCa
    d[:, :, 2] = ls
    return d


def oracle_trajectory(topo, p, z, pCa, n_steps, seed, mode="exact"):
    """Oracle states after 0..n_steps steps with the device's Philox draws."""
    cfg = mo.SolverConfig(mode=mode)
    pre = mo.preconditioner(topo, p)
    st = mo.init_state(topo, p, z, cfg, pre)
    st.ls = 14.0
    tab = mo.make_tables(p, pCa, 14.0, 1.0, topo.xb_bin_centers)
    S, X = topo.n_thin * topo.n_sites, topo.total_xbs
    states = [st.copy()]
    for t in range(n_steps):
        st, _ = mo.step(st, topo, p, tab, z, pCa, 14.0, 0.0, uniforms(seed, t, PURPOSE_TM, S),
                        uniforms(seed, t, PURPOSE_XB, X), cfg, pre)
        states.append(st.copy())
    return states, tab
