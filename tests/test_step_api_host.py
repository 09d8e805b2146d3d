"""Host side of the single-step API (multifil_jax_amd.timestep): the
State <-> engine record conversion, the NaN-sentinel resolution and the
Philox key successor.  No GPU calls."""
import numpy as np

from oracle import mf_oracle as mo


def _api_state(st, topo):
    from multifil_jax_amd.timestep import State, ThickState, ThinState
    shp = (topo.n_thick, topo.n_crowns, topo.n_xb_per_crown)
    return State(ThickState(st.thick.copy(), st.xb_state.reshape(shp), st.xb_bound.reshape(shp),
                            st.xb_nearest.reshape(shp),
                            np.stack([st.xb_xdist, np.full(st.xb_xdist.shape, 14.0, np.float32)], -1).reshape(shp + (2,))),
                 ThinState(st.thin.copy(), st.tm.copy(), st.thin_bound.copy()))


def test_record_round_trip_and_key(topo22, pdict):
    from multifil_jax_amd.timestep import _to_record, _from_record
    from gpu_helpers import oracle_trajectory
    states, _ = oracle_trajectory(topo22, pdict, 1000.0, 4.5, 5, seed=3)
    st = states[5]
    api = _api_state(st, topo22)
    rec = _to_record(api, topo22, 14.0)
    # the previous step's table key, as the oracle forms it (transitions.py:873-922)
    assert np.array_equal(rec["xb_key"], mo.xb_keys(st.xb_xdist, st.xb_nearest, st.tm, topo22))
    back = _from_record(rec, topo22, 14.0)
    for a, b in zip(back.thick + back.thin, api.thick + api.thin):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert back.thick.axial.dtype == np.float64


def test_realize_state_matches_the_oracle(topo22):
    from multifil_jax_amd.timestep import realize_state
    s = realize_state(topo22, 1000.0)
    o = mo.realize_state(topo22, 1000.0)
    assert np.array_equal(s.thick.axial, o.thick) and np.array_equal(s.thin.axial, o.thin)
    assert np.all(s.thick.xb_bound_to == -1) and np.all(s.thin.bound_to == -1)
    assert np.all(s.thick.xb_states == 0) and np.all(s.thin.tm_states == 0)


def test_resolve_value_and_keys():
    from multifil_jax_amd.timestep import resolve_value, split_key
    assert resolve_value(np.nan, 4.5) == np.float32(4.5)
    assert resolve_value(5.0, 4.5) == np.float32(5.0)
    keys = [np.uint64(7)]
    for _ in range(1000):
        keys.append(split_key(keys[-1]))
    assert len(set(int(k) for k in keys)) == len(keys)
    # SplitMix64 of state 0: the first output of the published generator
    assert split_key(0) == np.uint64(0xE220A8397B1DCDAF)
