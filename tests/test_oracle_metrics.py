"""Pin the oracle's 52-metric function (metrics_fn.py:79-417) by checks that
do not restate its formulas:

* the overlap band (metrics_fn.py:79-151) against a brute-force per-crown
  reach test: a site counts iff some crown of some thick filament lies
  within the 13 nm head reach of it (geometry.py:224) and the site is past
  the M-line; with crowns 14.3 nm apart the union of reach windows is the
  band [min crown - 13, max crown + 13];
* the energies against the independently written lattice energy whose
  gradient pins the residual (tests/test_oracle_mechanics.py): the titin
  energy is that energy's titin term (per titin, less its constant), and
  the first-spring energy changes with the first crown at the rate of that
  filament's M-line force;
* the expected-ATP / tear sums against one scipy.linalg.expm per head of the
  absorbing generator (rows 0 and 4 zeroed, transitions.py:1077) built at
  that head's own bin centre and permissiveness, not through the key table;
* counts, displacement statistics and work against plain loops.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import mf_oracle as mo

from gpu_helpers import oracle_trajectory
from test_oracle_mechanics import _energy_torch

M = {n: i for i, n in enumerate(mo.METRIC_NAMES)}


def _step_pair(topo, p, z=1000.0, pCa=4.5, k=10, seed=21):
    states, tab = oracle_trajectory(topo, p, z, pCa, k, seed)
    old = states[k - 1]
    from oracle.philox import uniforms, PURPOSE_TM, PURPOSE_XB
    S, X = topo.n_thin * topo.n_sites, topo.total_xbs
    pre = mo.preconditioner(topo, p)
    new, mt = mo.step(old, topo, p, tab, z, pCa, 14.0, 0.0, uniforms(seed, k, PURPOSE_TM, S),
                      uniforms(seed, k, PURPOSE_XB, X), mo.SolverConfig(mode="exact"), pre)
    return old, new, mt, tab


@pytest.mark.parametrize("z", [1000.0, 760.0])
def test_overlap_band_brute_force(topo22, pdict, z):
    old, new, mt, _ = _step_pair(topo22, pdict, z=z, k=4)
    crowns = topo22.crown_offsets.astype(np.float32).ravel()
    ns = topo22.n_sites
    n_ov = n2 = n3 = 0
    for i in range(topo22.n_thin):
        for s in range(ns):
            x = new.thin[i, s]
            reach = any(abs(x - c) <= 13.0 for c in crowns)
            if reach and x > 0.0:
                n_ov += 1
                n2 += new.tm[i, s] == 2
                n3 += new.tm[i, s] == 3
    assert mt['n_overlap_sites'] == n_ov
    assert n_ov < topo22.n_thin * ns                   # some sites lie outside the band
    assert mt['frac_tm_state_2_overlap'] == pytest.approx(n2 / n_ov, abs=1e-15)
    assert mt['frac_tm_state_3_overlap'] == pytest.approx(n3 / n_ov, abs=1e-15)
    assert mt['frac_tm_available_overlap'] == pytest.approx((n2 + n3) / n_ov, abs=1e-15)
    if z < 800:
        assert (new.thin <= 0).any()                    # the M-line mask is exercised


def test_titin_energy_is_the_energy_term(topo22, pdict):
    old, new, mt, _ = _step_pair(topo22, pdict)
    z, d = 1000.0, 14.0
    x = torch.tensor(new.x(), dtype=torch.float64)
    no_titin = dict(pdict, titin_a=0.0)
    e_tit = float(_energy_torch(topo22, pdict, z, d, new, x) - _energy_torch(topo22, no_titin, z, d, new, x))
    per_titin = e_tit / (6 * topo22.n_thick) - pdict['titin_a'] / pdict['titin_b']
    assert mt['titin_energy_avg'] == pytest.approx(per_titin, rel=1e-9)
    xo = torch.tensor(old.x(), dtype=torch.float64)
    e_old = float(_energy_torch(topo22, pdict, z, d, old, xo) - _energy_torch(topo22, no_titin, z, d, old, xo))
    assert mt['titin_energy_delta_avg'] == pytest.approx((e_tit - e_old) / (6 * topo22.n_thick), rel=1e-7, abs=1e-9)


def test_first_spring_energy_rate_is_the_mline_force(topo22, pdict):
    old, new, mt, _ = _step_pair(topo22, pdict)
    h = 1e-4
    for t in range(topo22.n_thick):
        vals = []
        for s in (+1, -1):
            n2 = new.copy()
            n2.thick[t, 0] += s * h
            m2 = mo.metrics(old, n2, topo22, pdict, 1000.0, 4.5, 14.0, old.thick, 0.0, 0.0, 0,
                            np.zeros((400, 6, 6)), np.zeros(topo22.total_xbs, np.int64))
            vals.append(m2['thick_energy_first_avg'] * topo22.n_thick)
        rate = (vals[0] - vals[1]) / (2 * h)
        one = np.zeros_like(new.thick)
        one[t] = new.thick[t]
        f_t = mo.axial_force(new.thick, topo22, pdict['thick_k']) - mo.axial_force(
            np.where(np.arange(topo22.n_thick)[:, None] == t, topo22.crown_offsets, new.thick), topo22, pdict['thick_k'])
        assert rate == pytest.approx(f_t, rel=1e-6, abs=1e-6)


def test_expected_atp_and_tear_by_per_head_expm(topo22, pdict):
    old, new, mt, tab = _step_pair(topo22, pdict, k=12)
    ox = old.xb_state.astype(int)
    assert (ox == 3).sum() > 5 and (ox == 2).sum() > 5
    edges = topo22.xb_bin_edges.astype(np.float32)
    centers = topo22.xb_bin_centers
    ns = topo22.n_sites
    atp = tear = 0.0
    for j in np.nonzero((ox == 2) | (ox == 3))[0]:
        b = int(np.searchsorted(edges, np.float32(old.xb_xdist[j]), side="right")) - 1
        b = min(max(b, 0), len(centers) - 1)
        site = min(max(int(old.xb_nearest[j]), 0), ns - 1)
        perm = float(old.tm[topo22.xb_to_thin_id[j], site] == 3 and topo22.xb_valid[j])
        Q = mo.xb_generators(pdict, 4.5, 14.0, np.array([centers[b]], np.float64), perm)[0]
        Q[0, :] = 0.0
        Q[4, :] = 0.0
        P = scipy.linalg.expm(Q * 1.0)
        if ox[j] == 3:
            atp += P[3, 4]
        tear += P[ox[j], 0]
    assert mt['atp_expected_p'] == pytest.approx(atp, rel=1e-7)
    assert mt['xb_tear_expected'] == pytest.approx(tear, rel=1e-7)
    assert mt['work_per_atp'] == pytest.approx(mt['work_thick'] / atp if atp > 0.01 else 0.0, rel=1e-7)


def test_counts_displacements_and_work_by_loops(topo22, pdict):
    old, new, mt, _ = _step_pair(topo22, pdict)
    X, S = topo22.total_xbs, topo22.n_thin * topo22.n_sites
    names = ('drx', 'loose', 'tight_1', 'tight_2', 'free_2', 'srx')
    for s, nm in enumerate(names):
        c = sum(1 for v in new.xb_state if v == s)
        assert mt['n_xb_' + nm] == c and mt['frac_xb_' + nm] == pytest.approx(c / X, abs=1e-15)
    assert mt['n_bound'] == sum(1 for v in new.xb_state if 1 <= v <= 3)
    for s in range(4):
        assert mt[f'n_tm_state_{s}'] == sum(1 for v in new.tm.ravel() if v == s)
    assert mt['actin_permissiveness'] == mt['frac_tm_state_3']
    assert mt['atp_consumed'] == sum(1 for a, b in zip(old.xb_state, new.xb_state) if a == 3 and b in (4, 0))
    assert mt['newly_bound'] == sum(1 for a, b in zip(old.xb_state, new.xb_state) if a == 0 and b == 1)
    # displacements from the rest positions: crowns at their offsets, sites at
    # the running sum of their rests along the filament
    dk = [new.thick[t, c] - float(np.float32(topo22.crown_offsets[t, c]))
          for t in range(topo22.n_thick) for c in range(topo22.n_crowns)]
    dn = []
    for i in range(topo22.n_thin):
        acc = np.float32(0.0)
        for s in range(topo22.n_sites):
            acc = np.float32(acc + np.float32(topo22.binding_rests[i, s]))
            dn.append(new.thin[i, s] - float(acc))
    for nm, a in (('thick', dk), ('thin', dn)):
        mean = sum(a) / len(a)
        var = sum((v - mean) ** 2 for v in a) / len(a)          # population std, like jnp.std
        assert mt[nm + '_displace_mean'] == pytest.approx(mean, abs=1e-9)
        assert mt[nm + '_displace_max'] == pytest.approx(max(a), abs=1e-12)
        assert mt[nm + '_displace_min'] == pytest.approx(min(a), abs=1e-12)
        assert mt[nm + '_displace_std'] == pytest.approx(var ** 0.5, rel=1e-9)
    pre = old.thick
    mean_dx = sum(new.thick[t, c] - pre[t, c] for t in range(topo22.n_thick)
                  for c in range(topo22.n_crowns)) / (topo22.n_thick * topo22.n_crowns)
    assert mt['work_thick'] == pytest.approx(mt['axial_force'] * mean_dx, rel=1e-9)
    assert mt['work_thick_mean'] == pytest.approx(mt['work_thick'] / (topo22.n_thick * topo22.n_crowns), rel=1e-12)
