"""Pin the oracle's kinetics by brute force: nearest sites (geometry.py:
60-262), TM neighbour configurations (transitions.py:384-431), inverse-CDF
sampling (transitions.py:618-619) and the binding bookkeeping
(transitions.py:1173-1230)."""
import numpy as np

from oracle import mf_oracle as mo


def test_nearest_sites_brute_force(topo22, rng):
    st = mo.realize_state(topo22, 1100.0)
    thick = (st.thick + rng.normal(0, 2, st.thick.shape)).astype(np.float32)
    thin = (st.thin + rng.normal(0, 2, st.thin.shape)).astype(np.float32)
    near, xd = mo.nearest_sites(thick, thin, topo22)
    t = topo22
    for j in range(t.total_xbs):
        crown = thick.ravel()[j // 3]
        head = np.float32(crown + np.float32(13.0))
        best, bj = np.inf, None
        sites = t.face_to_sites[t.xb_to_thin_id[j], t.xb_to_thin_face[j]]
        for s in sites:                       # candidate order; first minimum wins
            pos = thin[t.xb_to_thin_id[j], s]
            if pos <= 0:
                continue
            dd = abs(np.float32(pos - head))
            if dd < best:
                best, bj = dd, s
        if bj is None:
            bj = sites[0]
        assert near[j] == bj
        assert xd[j] == np.float32(thin[t.xb_to_thin_id[j], bj] - crown)


def test_tm_configs_brute_force(topo22, rng):
    t = topo22
    tm = rng.integers(0, 4, (t.n_thin, t.n_sites)).astype(np.int8)
    cfg = mo.tm_configs(tm, t).reshape(t.n_thin, t.n_sites)
    for i in range(t.n_thin):
        for s in range(t.n_sites):
            n2 = n3 = nc = 0
            for nb in (t.tm_prev_neighbor[i, s], t.tm_next_neighbor[i, s]):
                if nb == s:
                    continue
                v = tm[i, nb]
                n2 += v == 2
                n3 += v == 3
                nc += v <= 1
            assert cfg[i, s] == 9 * n2 + 3 * n3 + nc


def test_sampling_rule(rng):
    prob = rng.dirichlet(np.ones(6), 500)
    u = rng.random(500)
    got = mo.sample_rows(prob, u)
    for k in range(500):
        c = 0.0
        want = 0
        for j in range(6):
            c += prob[k, j]
            if u[k] < c:
                want = j
                break
        assert got[k] == want
    # no entry above u -> state 0 (transitions.py:618-619)
    assert mo.sample_rows(np.array([[0.2, 0.3, 0.49999]]), np.array([0.999995]))[0] == 0


def test_binding_bookkeeping_rules(topo22, pdict):
    t = topo22
    st = mo.realize_state(t, 1000.0)
    st.xb_nearest, st.xb_xdist = mo.nearest_sites(st.thick.astype(np.float32),
                                                  st.thin.astype(np.float32), t)
    tm = np.full((t.n_thin, t.n_sites), 3, np.int8)      # everything permissive
    # P: every head goes 0 -> 1 with certainty
    nb = t.n_xb_bins
    P = np.zeros((2 * nb, 6, 6))
    P[:, :, 0] = 1.0
    P[:, 0, 0], P[:, 0, 1] = 0.0, 1.0
    u = np.full(t.total_xbs, 0.5)
    xs, xb, tb = mo.xb_transition(st, tm, P, u, t)
    # heads sharing a nearest site: the lowest index owns it, all stay bound
    site = t.xb_to_thin_id * t.n_sites + st.xb_nearest
    for s_ in np.unique(site):
        js = np.nonzero(site == s_)[0]
        assert tb.ravel()[s_] == js.min()
        assert np.all(xs[js] == 1) and np.all(xb[js] == st.xb_nearest[js])
    # an occupied nearest site blocks binding and the head falls back to 0
    st2 = st.copy()
    st2.thin_bound = tb
    st2.xb_state = np.zeros_like(xs)
    st2.xb_bound = np.full_like(xb, -1)
    xs2, xb2, tb2 = mo.xb_transition(st2, tm, P, u, t)
    assert np.all(xs2 == 0) and np.all(xb2 == -1) and np.array_equal(tb2, tb)
    # unbinding clears the site
    P2 = np.zeros((2 * nb, 6, 6))
    P2[:, :, 0] = 1.0
    xs3, xb3, tb3 = mo.xb_transition(mo.State(st.thick, st.thin, xs, xb, st.xb_nearest, st.xb_xdist,
                                              tm, tb, 14.0), tm, P2, u, t)
    assert np.all(xs3 == 0) and np.all(xb3 == -1) and np.all(tb3 == -1)


def test_tm_lock(topo22, pdict):
    t = topo22
    tm = np.full((t.n_thin, t.n_sites), 3, np.int8)
    bound = np.full((t.n_thin, t.n_sites), -1, np.int32)
    bound[0, 5] = 17
    P = mo.expm_ref(mo.tm_generators(pdict, 9.0) * 50.0)     # strongly relaxing
    u = np.full(t.n_thin * t.n_sites, 0.999)
    new = mo.tm_transition(tm, bound, P, u, t)
    assert new[0, 5] == 3
