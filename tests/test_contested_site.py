"""The contested-site scenario of the bookkeeping test, checked on the CPU:
the oracle, fed the constructed state and draws, unbinds C from the site
and binds D to it, so the site ends owned (transitions.py:1203-1219)."""
import numpy as np

from oracle import mf_oracle as mo

from gpu_helpers import contested_site_state, oracle_trajectory


def test_contested_site_scenario(topo22, pdict):
    states, tab = oracle_trajectory(topo22, pdict, 1000.0, 4.5, 6, seed=11)
    st, u_tm, u_xb, C, D, th, site = contested_site_state(topo22, pdict, tab, states[6])
    assert st.thin_bound[th, site] == -1 and st.xb_bound[C] == site
    new, _ = mo.step(st, topo22, pdict, tab, 1000.0, 4.5, 14.0, 0.0, u_tm, u_xb,
                     mo.SolverConfig(mode="exact"), mo.preconditioner(topo22, pdict))
    assert new.xb_bound[C] == -1 and new.xb_state[C] in (0, 4, 5)
    assert new.xb_bound[D] == site and new.xb_state[D] == 1
    assert new.thin_bound[th, site] == min(D, *[j for j in np.nonzero(new.xb_bound == site)[0]
                                                 if topo22.xb_to_thin_id[j] == th])
