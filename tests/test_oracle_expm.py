"""Pin the oracle's matrix exponential (transitions.py:99-190) against scipy,
closed forms, semigroup and stochasticity."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import mf_oracle as mo


def _random_generators(rng, n, batch, scale):
    Q = rng.uniform(0, 1, (batch, n, n)) * scale[:, None, None]
    Q[:, np.arange(n), np.arange(n)] = 0
    Q[:, np.arange(n), np.arange(n)] = -Q.sum(axis=2)
    return Q


def test_expm_matches_scipy_random(rng):
    for n in (2, 4, 6):
        Q = _random_generators(rng, n, 200, 10.0 ** rng.uniform(-3, 4, 200))
        P = mo.expm_ref(Q)
        ref = np.array([sla.expm(q) for q in Q])
        assert np.max(np.abs(P - ref)) < 1e-8


def test_expm_model_tables_match_scipy(topo22, pdict):
    for pCa in (4.0, 5.5, 7.0):
        for d in (12.0, 14.0, 16.0):
            Q = mo.xb_bin_generators(pdict, pCa, d, topo22.xb_bin_centers)
            for A in (Q, mo.absorbing(Q)):
                P = mo.expm_ref(A)
                ref = np.array([sla.expm(q) for q in A])
                assert np.max(np.abs(P - ref)) < 1e-9
        Qt = mo.tm_generators(pdict, pCa)
        assert np.max(np.abs(mo.expm_ref(Qt) - np.array([sla.expm(q) for q in Qt]))) < 1e-10


def test_two_state_closed_form():
    # p(t) = a/(a+b) (1 - e^{-(a+b) t}) for 0 -> 1 with rates a, b.  The
    # reference's extra b6*A^6 term limits agreement to ~1e-10.
    for a, b, t in [(0.3, 0.7, 1.0), (1e3, 2.0, 1.0), (5.0, 5e3, 0.1), (1e-3, 1e-4, 1.0)]:
        Q = np.array([[[-a, a], [b, -b]]]) * t
        P = mo.expm_ref(Q)[0]
        want = a / (a + b) * (1 - np.exp(-(a + b) * t))
        assert abs(P[0, 1] - want) < 1e-10 + 1e-9 * want


def test_stochastic_and_semigroup(topo22, pdict):
    Q = mo.xb_bin_generators(pdict, 5.0, 14.0, topo22.xb_bin_centers)
    P1 = mo.expm_ref(Q * 1.0)
    P2 = mo.expm_ref(Q * 2.0)
    assert np.allclose(P1.sum(axis=2), 1.0, atol=1e-14)
    assert np.all(P1 >= -1e-15)
    assert np.max(np.abs(P1 @ P1 - P2)) < 1e-9


def test_scaling_power_rule(rng):
    # s = max(0, ceil(log2(||A||_inf / 0.5 + 1e-10))); brute force: smallest s
    # with ||A||/2^s <= 0.5 for norms away from powers of two
    norms = 10.0 ** rng.uniform(-4, 5, 500)
    A = np.zeros((500, 2, 2))
    A[:, 0, 0] = -norms / 2
    A[:, 0, 1] = norms / 2
    s = mo.scaling_power(A)
    for nrm, si in zip(norms, s):
        brute = 0
        while np.float32(nrm) / 2.0 ** brute > 0.5:
            brute += 1
        assert si == brute


def test_squaring_cap_is_18():
    # ||A|| beyond 2^17 is under-squared (the reference loops 18 times)
    A = np.array([[[-1.0, 1.0], [1.0, -1.0]]]) * 2.0 ** 20
    assert mo.scaling_power(A)[0] == 22
    P = mo.expm_ref(A)
    assert np.all(np.isfinite(P)) and np.allclose(P.sum(axis=2), 1.0)


def test_pade_solve_needs_no_pivoting(topo22, pdict):
    # The device solves (V-U) X = (V+U) without pivoting; partial pivoting
    # would never swap because V-U stays diagonally dominant by columns.
    Q = mo.xb_bin_generators(pdict, 4.5, 14.0, topo22.xb_bin_centers)
    for A in (Q, mo.absorbing(Q), mo.tm_generators(pdict, 4.5)):
        s = mo.scaling_power(A)
        As = A / np.power(2.0, s)[:, None, None]
        n = A.shape[-1]
        eye = np.eye(n)
        A2 = As @ As
        A4 = A2 @ A2
        b = mo.PADE_B
        U = As @ (b[1] * eye + b[3] * A2 + b[5] * A4)
        V = b[0] * eye + b[2] * A2 + b[4] * A4 + b[6] * A4 @ A2
        M = V - U
        for k in range(n):
            piv = np.argmax(np.abs(M[:, k:, k]), axis=1)
            assert np.all(piv == 0)
            f = M[:, k + 1:, k] / M[:, k, k][:, None]
            M[:, k + 1:, :] -= f[:, :, None] * M[:, k, None, :]


def test_absorbing_chain_from_its_transient_block(topo22, pdict):
    """The metrics' absorbing chain (states 0 and 4 absorbing,
    transitions.py:1029-1078) needs P[3][4], P[2][0] and P[3][0] of
    exp(Q_abs dt).  The device computes them from the 3x3 transient block:
    with A = T dt (rates among states 1-3) and exits r10 (1 -> 0), r34
    (3 -> 4), the absorption probabilities within dt are
    A^-1 (e^A - I) (R dt).  Checked here against scipy's 6x6 exponential of
    the absorbing generator over every bin, both permissiveness levels,
    three spacings and two calcium levels (a sign or an index slip in the
    block fails by orders of magnitude)."""
    c = np.asarray(topo22.xb_bin_centers, np.float64)
    worst = 0.0
    for pCa in (4.5, 6.5):
        for d in (11.0, 14.0, 16.5):
            for perm in (0.0, 1.0):
                Q = mo.absorbing(mo.xb_generators(pdict, pCa, d, c, perm=perm))
                for q in Q:
                    P = sla.expm(q * 1.0)
                    A = q[1:4, 1:4] * 1.0
                    e10, e34 = q[1, 0] * 1.0, q[3, 4] * 1.0
                    E = sla.expm(A)
                    y = np.linalg.solve(A, (E[:, 0] - np.eye(3)[:, 0]) * e10)
                    z = np.linalg.solve(A, (E[:, 2] - np.eye(3)[:, 2]) * e34)
                    got = np.array([z[2], y[1], y[2]])
                    want = np.array([P[3, 4], P[2, 0], P[3, 0]])
                    assert np.all(got >= -1e-15)
                    worst = max(worst, float(np.max(np.abs(got - want))))
                    assert np.allclose(got, want, rtol=1e-9, atol=1e-15), (pCa, d, perm, got, want)
    assert worst < 1e-12
