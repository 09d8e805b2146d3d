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
                xb_key=key.astype(np.int32), xb_near=st.xb_nearest.astype(np.int32),
                xb_xdist=st.xb_xdist.astype(np.float32), thin_bound=st.thin_bound.ravel().astype(np.int32),
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


def contested_site_state(topo, p, tab, st, z=1000.0):
    """A state in which, in one step, a head C unbinds from a site that the
    site's owner record already shows as free (thin_bound = -1, left so by
    an earlier unbinding while C stayed bound), and a head D binds that
    same site (transitions.py:1203-1219: clears land before binds).

    Built from an oracle state ``st`` by editing C, D and the site; returns
    (state, u_tm, u_xb, C, D, thin, site).  The uniforms of C, D and the
    site are picked from the oracle's own table rows; every other draw is
    the Philox stream of seed 77.
    """
    ns, nb = topo.n_sites, topo.n_xb_bins
    S, X = topo.n_thin * ns, topo.total_xbs
    s0 = st.copy()
    near, xd = mo.nearest_sites(s0.thick.astype(np.float32), s0.thin.astype(np.float32), topo)
    tid = topo.xb_to_thin_id
    u_tm = uniforms(77, 0, PURPOSE_TM, S).astype(np.float64)
    u_xb = uniforms(77, 0, PURPOSE_XB, X).astype(np.float64)
    # D: an unbound, target-valid head whose site is free, with the best
    # binding probability once that site is open (state 3)
    bins = mo.bin_index(xd, topo.xb_bin_edges)
    p01 = tab.P_xb[nb + bins, 0, 1]
    occ = s0.thin_bound.reshape(-1)[tid * ns + near] >= 0
    ok = (s0.xb_state == 0) & topo.xb_valid.astype(bool) & ~occ
    D = int(np.argmax(np.where(ok, p01, -1.0)))
    th, site = int(tid[D]), int(near[D])
    # C: another unbound head aimed at the same thin filament, the one whose
    # crown lies nearest the site, far from D in index (other waves)
    pos_s = s0.thin[th, site]
    crown = np.repeat(s0.thick.reshape(-1), topo.n_xb_per_crown)
    cand = (tid == th) & (s0.xb_state == 0) & (np.arange(X) != D) & (np.abs(np.arange(X) - D) > 64)
    C = int(np.argmin(np.where(cand, np.abs(pos_s - crown), np.inf)))
    s0.xb_state[C] = 1
    s0.xb_bound[C] = site
    s0.thin_bound[th, site] = -1
    s0.tm[th, site] = 3
    # the site stays open: u inside its 3 -> 3 interval
    cfg = mo.tm_configs(s0.tm, topo)[th * ns + site]
    c = np.cumsum(tab.P_tm[cfg, 3])
    u_tm[th * ns + site] = 0.5 * (c[2] + c[3])
    # D draws 0 -> 1; C draws 1 -> 0 (its row at its own nearest-site key)
    cD = np.cumsum(tab.P_xb[nb + bins[D], 0])
    u_xb[D] = 0.5 * (cD[0] + cD[1])
    tm_next = s0.tm.copy()
    keyC = mo.xb_keys(xd[C:C + 1], near[C:C + 1], tm_next, topo)[0]
    cC = np.cumsum(tab.P_xb[keyC, 1])
    u_xb[C] = 0.5 * cC[0]
    s0.xb_nearest, s0.xb_xdist = near, xd
    return s0, u_tm.astype(np.float32), u_xb.astype(np.float32), C, D, th, site
