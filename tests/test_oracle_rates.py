"""Pin the oracle's rate laws (rate_functions.py, transitions.py:434-513,
632-804) against identities and the values the reference states in its
docstrings."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import mf_oracle as mo


def _h(p, cfg):
    n2, n3, nc = cfg // 9, (cfg // 3) % 3, cfg % 3
    return p['tm_J_C'] * n2 + p['tm_J_M'] * n3 - 0.5 * (p['tm_J_C'] + p['tm_J_M']) * nc


def test_tm_detailed_balance_shift(pdict):
    # k01/k10 = Keq01 [Ca] e^h, k12/k21 = Keq12 e^h, k23/k32 = Keq23 e^h
    # (transitions.py:362-372); k30 unscaled.
    pCa = 5.3
    ca = 10.0 ** -pCa
    Q = mo.tm_generators(pdict, pCa)
    for cfg in range(27):
        eh = np.exp(_h(pdict, cfg))
        assert np.isclose(Q[cfg, 0, 1] / Q[cfg, 1, 0], pdict['tm_Keq_01'] * ca * eh, rtol=1e-12)
        assert np.isclose(Q[cfg, 1, 2] / Q[cfg, 2, 1], pdict['tm_Keq_12'] * eh, rtol=1e-12)
        assert np.isclose(Q[cfg, 2, 3] / Q[cfg, 3, 2], pdict['tm_Keq_23'] * eh, rtol=1e-12)
        assert Q[cfg, 3, 0] == pdict['tm_k_30']
        assert Q[cfg, 0, 2] == Q[cfg, 0, 3] == Q[cfg, 1, 3] == Q[cfg, 3, 1] == 0.0
    assert np.allclose(Q.sum(axis=2), 0.0, atol=1e-9)


def test_tm_uncoupled_limit(pdict):
    # J_C = J_M = 0 reduces to the uncoupled chain (transitions.py:374-376)
    p = dict(pdict, tm_J_C=0.0, tm_J_M=0.0)
    Q = mo.tm_generators(p, 6.0)
    assert np.allclose(Q, Q[0][None], rtol=0, atol=0)


def test_tm_config_index_order(pdict):
    # index = 9 n2 + 3 n3 + n_closed (meshgrid 'ij', transitions.py:464-468)
    p = dict(pdict, tm_J_C=0.7, tm_J_M=0.0)
    Q = mo.tm_generators(p, 6.0)
    # only n2 changes the field when J_M = 0 and J_C != 0 ... and n_closed
    assert np.isclose(Q[9, 1, 2] / Q[0, 1, 2], np.exp(0.35))


def test_srx_equilibria(pdict):
    # r05 / (r05 + r50): skeletal ~1/2 at rest (rate_functions.py:735-741);
    # cardiac 97.5 % at rest and ~33 % at pCa 4.5 (params.py:939-943)
    def srx(p, pCa):
        Q = mo.xb_generators(p, pCa, 14.0, np.array([5.0]), 0.0)
        r05, r50 = Q[0, 0, 5], Q[0, 5, 0]
        return r05 / (r05 + r50)
    assert abs(srx(pdict, 9.0) - 0.5) < 1e-6
    from multifil_jax_amd.params import get_cardiac_params, DYNAMIC_FIELDS
    _, dc = get_cardiac_params()
    pc = {n: float(getattr(dc, n)) for n in DYNAMIC_FIELDS}
    assert abs(srx(pc, 9.0) - 0.975) < 1e-3
    assert abs(srx(pc, 4.5) - 1 / 3) < 5e-3


def test_r32_over_r23_is_exp_minus_6(pdict):
    # r32 / r23 = exp(U_tight_2 - U_tight_1) = e^-6 below the caps
    # (rate_functions.py:530-533)
    x = np.linspace(-2.0, 12.0, 15)
    Q = mo.xb_generators(pdict, 4.5, 14.0, x, 1.0)
    r23, r32 = Q[:, 2, 3], Q[:, 3, 2]
    ok = r23 < 1e4 * np.exp(-6)
    assert ok.any()
    assert np.allclose(r32[ok] / r23[ok], np.exp(-6.0), rtol=1e-10)


def test_caps_and_permissiveness_gate(pdict):
    x = np.linspace(-8, 35, 50)
    Q0 = mo.xb_generators(pdict, 4.5, 14.0, x, 0.0)
    Q1 = mo.xb_generators(pdict, 4.5, 14.0, x, 1.0)
    assert np.all(Q0[:, 0, 1] == 0.0) and np.all(Q1[:, 0, 1] > 0.0)
    off = Q1.copy()
    off[:, np.arange(6), np.arange(6)] = 0
    assert off.max() <= 1e4 + 1e-9
    assert np.allclose(Q1.sum(axis=2), 0.0, atol=1e-9)
    assert np.all(Q1[:, 4, 3] == 0.0)   # r43 structurally zero


def _xb_axial_force(p, x, d, strong):
    m = mo._xb_springs(p, np.array(strong))
    gk, gr, ck, cr = [float(v) for v in m]
    r = np.hypot(x, d)
    th = np.arctan2(d, x)
    return gk * (r - gr) * x / r - (ck / r) * (th - cr) * d / r


def test_force_zero_points_and_stroke(pdict):
    # params.py:171-174: at d = 14 nm the axial force vanishes at x = 14.126
    # (weak) and 8.288 nm (strong); effective stroke 5.84 nm.
    xw = brentq(lambda x: _xb_axial_force(pdict, x, 14.0, False), 5, 25)
    xs = brentq(lambda x: _xb_axial_force(pdict, x, 14.0, True), 2, 14)
    assert abs(xw - 14.126) < 5e-4
    assert abs(xs - 8.288) < 5e-4
    assert abs((xw - xs) - 5.84) < 5e-3


def test_projected_strong_stiffness(pdict):
    # params.py:183-184,229-231: projected axial stiffness ~1.33 pN/nm at d=14
    xs = brentq(lambda x: _xb_axial_force(pdict, x, 14.0, True), 2, 14)
    h = 1e-5
    k = (_xb_axial_force(pdict, xs + h, 14.0, True) - _xb_axial_force(pdict, xs - h, 14.0, True)) / (2 * h)
    assert abs(k - 1.33) < 0.03


def test_kT_value():
    assert np.isclose(mo.kT_pNnm(26.15), 1.3810e-23 * 299.3 * 1e21)


# ---------------------------------------------------------------------------
# XB rate laws (rate_functions.py:318-753, transitions.py:632-804), pinned by
# invariants that tie each rate to another rate or to the force kernel's
# expression (forces.py:398-399, itself pinned by autograd and the zero
# points above): E_weak(x) - E_weak(x0) = int f_weak dx, and likewise for
# the strong configuration.  A dropped term, a flipped sign or a wrong
# factor in any one rate law breaks the identity it enters.
# ---------------------------------------------------------------------------
D0 = 14.0
XS = np.linspace(-6.0, 30.0, 37)


def _gen(p, x, d=D0, perm=1.0, pCa=4.5):
    Q = mo.xb_generators(p, pCa, d, np.asarray(x, float), perm)
    return {k: Q[:, i, j] for k, (i, j) in
            dict(r01=(0, 1), r10=(1, 0), r12=(1, 2), r21=(2, 1), r23=(2, 3), r32=(3, 2),
                 r34=(3, 4), r40=(4, 0), r04=(0, 4), r05=(0, 5), r50=(5, 0)).items()}


def _energy_change(p, x0, x, strong, d=D0):
    """(E(x) - E(x0)) / kT of one head configuration, from the axial force."""
    from scipy.integrate import quad
    kt = mo.kT_pNnm(p['temp_celsius'])
    return quad(lambda s: _xb_axial_force(p, s, d, strong), x0, x, epsabs=1e-12, epsrel=1e-12)[0] / kt


def test_r01_boltzmann_window(pdict):
    # r01 = perm A01 exp(-E_weak): log r01 changes by -dE_weak / kT
    g = _gen(pdict, XS)
    x0 = 10.0
    g0 = _gen(pdict, [x0])
    for x, r in zip(XS, g['r01']):
        assert np.isclose(np.log(r / g0['r01'][0]), -_energy_change(pdict, x0, x, False), rtol=1e-7, atol=1e-9)
    # the window peaks at the weak zero-force point (14.126 nm at d = 14)
    from scipy.optimize import minimize_scalar
    m = minimize_scalar(lambda x: -_gen(pdict, [x])['r01'][0], bounds=(5, 25), method='bounded',
                        options={'xatol': 1e-6})
    assert abs(m.x - 14.126) < 1e-3
    assert np.all(g['r01'] <= pdict['xb_r01_coeff'])
    assert np.all(_gen(pdict, XS, perm=0.0)['r01'] == 0.0)


def test_r10_detailed_balance(pdict):
    # r10 = (r01 + 0.005) e^(U_loose - U_DRX), U_loose = U_loose0 + E_weak, so
    # r10 / (r01 + 0.005) * r01 / A01 = e^(U_loose0 - U_DRX0) wherever r10 is
    # below its cap
    g = _gen(pdict, XS)
    ok = g['r10'] < 1e4
    assert ok.sum() > 10
    lhs = g['r10'][ok] / (g['r01'][ok] + 0.005) * g['r01'][ok] / pdict['xb_r01_coeff']
    assert np.allclose(lhs, np.exp(pdict['xb_U_loose'] - pdict['xb_U_DRX']), rtol=1e-9)
    # a covered site: r01 = 0, r10 = 0.005 e^(U_loose - U_DRX) still strain dependent
    g0 = _gen(pdict, XS, perm=0.0)
    kt = mo.kT_pNnm(pdict['temp_celsius'])
    ok0 = g0['r10'] < 1e4
    e = np.array([_energy_change(pdict, 10.0, x, False) for x in XS[ok0]])
    ref = 0.005 * np.exp(pdict['xb_U_loose'] - pdict['xb_U_DRX'])
    ew10 = -np.log(_gen(pdict, [10.0])['r01'][0] / pdict['xb_r01_coeff'])
    assert np.allclose(g0['r10'][ok0], ref * np.exp(ew10 + e), rtol=1e-7)
    assert kt > 0


def test_r12_half_energy_difference(pdict):
    # r12 = A12 e^(min(E_diff, 30) / 2), E_diff = E_weak - E_strong: log r12
    # changes by (dE_weak - dE_strong) / 2 kT, and r12 = A12 where E_diff = 0
    g = _gen(pdict, XS)
    x0 = 8.0
    g0 = _gen(pdict, [x0])
    for x, r in zip(XS, g['r12']):
        de = _energy_change(pdict, x0, x, False) - _energy_change(pdict, x0, x, True)
        ediff0 = 2.0 * np.log(g0['r12'][0] / pdict['xb_r12_coeff'])
        if ediff0 + de < 30.0:
            assert np.isclose(np.log(r / g0['r12'][0]), 0.5 * de, rtol=1e-7, atol=1e-9), x
        else:
            assert np.isclose(r, pdict['xb_r12_coeff'] * np.exp(15.0), rtol=1e-9)


def test_r21_detailed_balance(pdict):
    # r21 = r12 e^(U_t1 - U_loose) = r12 e^(U_t1_0 - U_loose0 - E_diff) and
    # r12 / A12 = e^(E_diff / 2): (r21 / r12) (r12 / A12)^2 = e^(U_t1_0 - U_loose0)
    g = _gen(pdict, XS)
    ediff = 2.0 * np.log(g['r12'] / pdict['xb_r12_coeff'])
    ok = (g['r21'] < 1e4) & (ediff < 30.0 - 1e-9)
    assert ok.sum() > 10
    lhs = g['r21'][ok] / g['r12'][ok] * (g['r12'][ok] / pdict['xb_r12_coeff']) ** 2
    assert np.allclose(lhs, np.exp(pdict['xb_U_tight_1'] - pdict['xb_U_loose']), rtol=1e-9)


def _f_strong(p, x, d=D0):
    # the rates' strong-state force, transitions.py:737 (not the force
    # kernel's expression): g_k (r - g_rest) + (1/r) c_k (theta - c_rest)
    r = np.hypot(x, d)
    th = np.arctan2(d, x)
    return p['xb_g_k_strong'] * (r - p['xb_g_rest_strong']) + p['xb_c_k_strong'] * (th - p['xb_c_rest_strong']) / r


def test_bell_and_slip_laws(pdict):
    # r23 = A23 e^(-f d23 / kT) (resisting load slows it), r34 = A34 e^(+f d34 / kT)
    kt = mo.kT_pNnm(pdict['temp_celsius'])
    g = _gen(pdict, XS)
    f = _f_strong(pdict, XS)
    ok = g['r23'] < 1e4
    assert np.allclose(g['r23'][ok] * np.exp(f[ok] * pdict['xb_delta_23'] / kt), pdict['xb_r23_coeff'], rtol=1e-9)
    ok = g['r34'] < 1e4
    assert ok.sum() > 10
    assert np.allclose(g['r34'][ok] * np.exp(-f[ok] * pdict['xb_delta_34'] / kt), pdict['xb_r34_coeff'], rtol=1e-9)
    # monotone in the load: more force, slower 2->3, faster 3->4
    order = np.argsort(f)
    assert np.all(np.diff(g['r23'][order]) <= 0) and np.all(np.diff(g['r34'][order]) >= 0)
    assert (g['r34'] == 1e4).any()          # the cap binds at large strain


def test_r32_spans_the_stated_range(pdict):
    # rate_functions.py:541-542: over x in [-4, 21.5] nm r32 varies ~8.5e5-fold
    r32 = _gen(pdict, np.linspace(-4.0, 21.5, 4001))['r32']
    assert 8.0e5 < r32.max() / r32.min() < 9.0e5


def test_state_independent_rates(pdict):
    # r40, r04, r05 constants; r50 a Hill function of [Ca] (rate_functions.py:726)
    for pCa in (9.0, 6.0, 4.5):
        g = _gen(pdict, XS, pCa=pCa)
        for k in ('r40', 'r04', 'r05', 'r50'):
            assert np.all(g[k] == g[k][0])
        ca, b = 10.0 ** -pCa, pdict['xb_srx_b']
        hill = pdict['xb_srx_k0'] + (pdict['xb_srx_kmax'] - pdict['xb_srx_k0']) * ca ** b / (
            pdict['xb_srx_ca50'] ** b + ca ** b)
        assert np.isclose(g['r50'][0], hill, rtol=1e-12)
        assert g['r40'][0] == pdict['xb_r40'] and g['r04'][0] == pdict['xb_r04']
