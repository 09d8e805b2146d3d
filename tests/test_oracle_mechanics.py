"""Pin the oracle's mechanics: residual vs the gradient of an independently
written energy (torch float64 autograd, the helper.py:145-342 pattern), JVP
vs central differences, the preconditioner vs closed-form Green's functions,
Newton-CG vs an exact sparse Newton solve, and the M-line force identity."""
import numpy as np
import pytest
import torch

from oracle import mf_oracle as mo


def _bound_state(topo, rng, frac=0.4, z=1000.0, pos_noise=0.3):
    st = mo.realize_state(topo, z)
    X = topo.total_xbs
    near, _ = mo.nearest_sites(st.thick.astype(np.float32), st.thin.astype(np.float32), topo)
    pick = rng.random(X) < frac
    # one XB per site at most, like the sampler without duplicates
    seen = set()
    for j in np.nonzero(pick)[0]:
        key = (topo.xb_to_thin_id[j], near[j])
        if key in seen:
            continue
        seen.add(key)
        st.xb_state[j] = rng.integers(1, 4)
        st.xb_bound[j] = near[j]
    st.thick = st.thick + rng.normal(0, pos_noise, st.thick.shape)
    st.thin = st.thin + rng.normal(0, pos_noise, st.thin.shape)
    return st


def _energy_torch(topo, p, z, d, st, x):
    """Total energy (pN nm) of the lattice, written from the model definition."""
    nk = topo.n_thick * topo.n_crowns
    xk = x[:nk].reshape(topo.n_thick, topo.n_crowns)
    xn = x[nk:].reshape(topo.n_thin, topo.n_sites)
    zero = torch.zeros((topo.n_thick, 1), dtype=x.dtype)
    ext_k = torch.diff(torch.cat([zero, xk], 1), dim=1) - torch.as_tensor(topo.crown_rests, dtype=x.dtype)
    E = 0.5 * p['thick_k'] * (ext_k ** 2).sum()
    zz = torch.full((topo.n_thin, 1), z, dtype=x.dtype)
    ext_n = torch.diff(torch.cat([xn, zz], 1), dim=1) - torch.as_tensor(topo.binding_rests, dtype=x.dtype)
    E = E + 0.5 * p['thin_k'] * (ext_n ** 2).sum()
    L = torch.sqrt((z - xk[:, -1]) ** 2 + d * d)
    E = E + 6 * (p['titin_a'] / p['titin_b']) * torch.exp(p['titin_b'] * (L - p['titin_rest'])).sum()
    s = st.xb_state.astype(int)
    bound = (s >= 1) & (s <= 3) & (st.xb_bound >= 0)
    for j in np.nonzero(bound)[0]:
        strong = s[j] in (2, 3)
        gk = p['xb_g_k_strong'] if strong else p['xb_g_k_weak']
        gr = p['xb_g_rest_strong'] if strong else p['xb_g_rest_weak']
        ck = p['xb_c_k_strong'] if strong else p['xb_c_k_weak']
        cr = p['xb_c_rest_strong'] if strong else p['xb_c_rest_weak']
        c = j // topo.n_xb_per_crown
        site = nk + topo.xb_to_thin_id[j] * topo.n_sites + st.xb_bound[j]
        xd = x[site] - x[c]
        r = torch.sqrt(xd * xd + d * d)
        th = torch.atan2(torch.as_tensor(d, dtype=x.dtype), xd)
        E = E + 0.5 * gk * (r - gr) ** 2 + 0.5 * ck * (th - cr) ** 2
    return E


def test_residual_is_minus_energy_gradient(topo22, pdict, rng):
    st = _bound_state(topo22, rng)
    z, d = 1000.0, 14.0
    m = mo.Mech(topo22, pdict, z, d, st.xb_state, st.xb_bound)
    x = st.x()
    F = mo.residual(m, x)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    E = _energy_torch(topo22, pdict, z, d, st, xt)
    (g,) = torch.autograd.grad(E, xt)
    assert m.bound.sum() > 50
    assert np.max(np.abs(F + g.numpy())) < 1e-7 * max(1.0, np.max(np.abs(F)))


def test_jvp_matches_central_differences(topo22, pdict, rng):
    st = _bound_state(topo22, rng)
    m = mo.Mech(topo22, pdict, 1000.0, 14.0, st.xb_state, st.xb_bound)
    x = st.x()
    kx = mo.xb_stiffness(m, x)
    kt = mo.titin_tip_stiffness(m, x)
    for _ in range(3):
        v = rng.normal(size=x.shape)
        h = 1e-5
        fd = (mo.residual(m, x + h * v) - mo.residual(m, x - h * v)) / (2 * h)
        jv = mo.jvp(m, kx, kt, v)
        assert np.max(np.abs(fd - jv)) < 1e-5 * np.max(np.abs(jv))
        J = mo.jacobian_sparse(m, x)
        assert np.max(np.abs(J @ v - jv)) < 1e-9 * np.max(np.abs(jv))


def test_preconditioner_is_exact_backbone_inverse(topo22, pdict, rng):
    apply = mo.preconditioner(topo22, pdict, eps=0.0)
    nc, ns = topo22.n_crowns, topo22.n_sites
    k, kn = pdict['thick_k'], pdict['thin_k']
    i = np.arange(nc)
    Gk = (np.minimum(i[:, None], i[None, :]) + 1) / k            # M-line anchored, tip free
    j = np.arange(ns)
    Gn = (ns - np.maximum(j[:, None], j[None, :])) / kn           # free end, Z-line anchored
    v = rng.normal(size=topo22.n_dof)
    got = apply(v)
    nk = topo22.n_thick * nc
    want = np.concatenate([(v[:nk].reshape(-1, nc) @ Gk).ravel(), (v[nk:].reshape(-1, ns) @ Gn).ravel()])
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    # and it inverts the negated passive Jacobian with no XBs attached
    st = mo.realize_state(topo22, 900.0)
    m = mo.Mech(topo22, pdict, 900.0, 14.0, st.xb_state, st.xb_bound)
    J = mo.jacobian_sparse(m, st.x())
    kt = mo.titin_tip_stiffness(m, st.x())
    Jb = J.toarray()
    tips = np.arange(topo22.n_thick) * nc + nc - 1
    Jb[tips, tips] -= kt                                        # titin is outside M
    assert np.max(np.abs(apply(-(Jb @ v)) - v)) < 1e-10 * np.max(np.abs(v))


def test_newton_cg_reaches_tolerance_and_equilibrium(topo22, pdict, rng):
    st = _bound_state(topo22, rng, pos_noise=0.05)
    m = mo.Mech(topo22, pdict, 1000.0, 14.0, st.xb_state, st.xb_bound)
    pre = mo.preconditioner(topo22, pdict)
    tol = mo.solver_tolerance(pdict)
    assert tol == 0.75
    x1, it, res = mo.newton_cg(m, st.x(), pre, tol, 4, 6)
    assert res <= tol and 1 <= it <= 4
    # FP64 floor: ~|x| eps k = 1000 * 2.2e-16 * 7500 ~ 2e-9 pN
    x2, it2, res2 = mo.newton_exact(m, st.x(), 1e-8)
    assert res2 < 1e-8
    # positions agree to the tolerance band of the reference's stopping rule
    assert np.max(np.abs(x1 - x2)) < 0.05
    # M-line force identity: per thick, k (x0 - bare) = titin + sum of XB forces
    F = mo.residual(m, x2)
    force = mo.axial_force(x2.reshape(-1)[:m.nk].reshape(topo22.n_thick, -1), topo22, pdict['thick_k'])
    xd, r, th = mo.xb_geometry(m, x2, 14.0)
    gk, gr, ck, cr = mo._xb_springs(pdict, m.strong)
    f = np.where(m.bound, gk * (r - gr) * xd / r - (ck / r) * (th - cr) * 14.0 / r, 0.0)
    xk = x2[:m.nk].reshape(topo22.n_thick, -1)
    ad = 1000.0 - xk[:, -1]
    L = np.hypot(ad, 14.0)
    tt = 6 * pdict['titin_a'] * np.exp(pdict['titin_b'] * (L - pdict['titin_rest'])) * ad / L
    assert abs(force - (tt.sum() + f.sum())) < 1e-6 * abs(force)


def test_dynamic_bordered_solve(topo22, pdict, rng):
    st = _bound_state(topo22, rng, pos_noise=0.05)
    m = mo.Mech(topo22, pdict, 1000.0, 14.0, st.xb_state, st.xb_bound)
    K = 5.0 * topo22.n_thick
    x, d, it, res = mo.newton_exact_dynamic(m, st.x(), 14.0, K, 14.0)
    assert res < 1e-8 and d > 1.0
    # bordered partials vs finite differences
    x0 = st.x()
    fax_d, frad_x, frad_d = mo.aug_partials(m, x0, 14.3)
    h = 1e-6
    fd = (mo.residual(m, x0, 14.3 + h) - mo.residual(m, x0, 14.3 - h)) / (2 * h)
    assert np.max(np.abs(fd - fax_d)) < 1e-6 * max(1, np.max(np.abs(fax_d)))
    g = lambda dd: mo.radial_residual(m, x0, dd, 0.0, 0.0)
    assert abs((g(14.3 + h) - g(14.3 - h)) / (2 * h) - frad_d) < 1e-5 * max(1, abs(frad_d))
    v = rng.normal(size=x0.shape)
    gx = (mo.radial_residual(m, x0 + h * v, 14.3, 0, 0) - mo.radial_residual(m, x0 - h * v, 14.3, 0, 0)) / (2 * h)
    assert abs(gx - frad_x @ v) < 1e-5 * max(1, abs(gx))
    # the CG variant lands in the same tolerance band
    pre = mo.preconditioner(topo22, pdict)
    x1, d1, it1, res1 = mo.newton_cg_dynamic(m, st.x(), 14.0, K, 14.0, pre, 0.75, 4, 6)
    assert res1 <= 0.75 and abs(d1 - d) < 1e-3
