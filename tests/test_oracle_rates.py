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
