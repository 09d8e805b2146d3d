"""FP64 NumPy restatement of the reference's per-timestep hot path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the checker for the
device kernels and the source of the CPU baseline, never product code.

Every function follows one reference function, cited as file:line into the
reference repository (multifil_jax/...).  Arithmetic is float64, except where
the reference takes a discrete decision from float32 values (nearest-site
argmin, bin digitize, M-line mask): those are evaluated on float32 copies in
float32, so the device and the oracle take the same decisions.

Conventions shared with the device path
---------------------------------------
* DOF vector x = [thick (Tk*nc) ; thin (Tn*ns)] (solver.py:727).
* Uniforms are injected (``u_tm`` of length S, ``u_xb`` of length X), the
  reference's ``random_values=`` seam (transitions.py:521,1086).
* Duplicate-index scatters in the binding bookkeeping (transitions.py:
  1205-1219) are resolved deterministically: clears and binds take precedence
  over the no-op rewrites, and when several heads bind one free site in the
  same step the lowest XB index owns it.  (XLA leaves this order undefined.)

A topology is any object with the reference's SarcTopology field names; a
parameter set is a mapping from DynamicParams field names to floats.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, Mapping, Optional, Tuple

import numpy as np

# Pade coefficients exactly as the reference lists them (transitions.py:84-92):
# b0..b5 are the [5/5] Pade coefficients, b6 an extra even term.
PADE_B = (1.0, 1.0 / 2.0, 1.0 / 9.0, 1.0 / 72.0, 1.0 / 1008.0, 1.0 / 30240.0, 1.0 / 1209600.0)
N_SQUARINGS_CAP = 18          # transitions.py:176-181
RATE_CAP = 10000.0            # rate_functions.py:384-624
MIN_FLOAT32_TOLERANCE = 0.25  # solver.py:105
N_TITIN_PER_THICK = 6         # forces.py:90,646
XB_REACH = np.float32(13.0)   # geometry.py:224

METRIC_NAMES = (
    'axial_force', 'solver_residual', 'z_line', 'pCa', 'lattice_spacing', 'n_bound',
    'n_xb_drx', 'n_xb_loose', 'n_xb_tight_1', 'n_xb_tight_2', 'n_xb_free_2', 'n_xb_srx',
    'frac_xb_bound', 'frac_xb_drx', 'frac_xb_loose', 'frac_xb_tight_1', 'frac_xb_tight_2',
    'frac_xb_free_2', 'frac_xb_srx', 'n_tm_state_0', 'n_tm_state_1', 'n_tm_state_2',
    'n_tm_state_3', 'frac_tm_state_0', 'frac_tm_state_1', 'frac_tm_state_2',
    'frac_tm_state_3', 'actin_permissiveness', 'frac_tm_state_2_overlap',
    'frac_tm_state_3_overlap', 'frac_tm_available_overlap', 'n_overlap_sites',
    'atp_consumed', 'newly_bound', 'thick_displace_mean', 'thick_displace_max',
    'thick_displace_min', 'thick_displace_std', 'thin_displace_mean', 'thin_displace_max',
    'thin_displace_min', 'thin_displace_std', 'thick_energy_first_avg',
    'thick_energy_first_delta_avg', 'titin_energy_avg', 'titin_energy_delta_avg',
    'work_thick', 'work_thick_mean', 'atp_expected_p', 'xb_tear_expected', 'work_per_atp',
    'newton_iters')


# ============================================================================
# matrix exponential
# ============================================================================

def scaling_power(A: np.ndarray) -> np.ndarray:
    """s = max(0, ceil(log2(||A||_inf / 0.5 + 1e-10))) (transitions.py:145-149).

    Evaluated in float32 like the reference, because s is a discrete choice.
    """
    norms = np.max(np.sum(np.abs(A), axis=-1), axis=-1).astype(np.float32)
    with np.errstate(divide='ignore'):
        lg = np.log2((norms / np.float32(0.5) + np.float32(1e-10)).astype(np.float32))
    return np.maximum(0, np.ceil(lg).astype(np.int64))


def expm_ref(A: np.ndarray) -> np.ndarray:
    """Batched scaling-and-squaring expm, the reference's variant.

    transitions.py:99-190: Pade-type approximant with PADE_B, solve
    (V-U) R = (V+U), square s times (at most 18), renormalise rows, NaN -> 1/n.
    A: (batch, n, n) float64.
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[-1]
    s = scaling_power(A)
    As = A / np.power(2.0, s)[:, None, None]
    eye = np.eye(n)
    A2 = As @ As
    A4 = A2 @ A2
    A6 = A4 @ A2
    b = PADE_B
    U = As @ (b[1] * eye + b[3] * A2 + b[5] * A4)
    V = b[0] * eye + b[2] * A2 + b[4] * A4 + b[6] * A6
    R = np.linalg.solve(V - U, V + U)
    for i in range(N_SQUARINGS_CAP):
        sq = R @ R
        R = np.where((i < s)[:, None, None], sq, R)
    with np.errstate(invalid='ignore', divide='ignore'):
        R = R / np.sum(R, axis=-1, keepdims=True)
    return np.where(np.isnan(R), 1.0 / n, R)


# ============================================================================
# rate matrices
# ============================================================================

def tm_generators(p: Mapping[str, float], pCa: float) -> np.ndarray:
    """27 TM generators indexed [9*n2 + 3*n3 + n_closed] (transitions.py:434-513).

    h = J_C n2 + J_M n3 - (J_C+J_M)/2 n_closed; forward rates x e^{h/2},
    reverse rates x e^{-h/2}, k_30 unscaled (rate_functions.py:165-311).
    """
    ca = 10.0 ** (-float(pCa))
    lv = np.arange(3, dtype=np.float64)
    n2, n3, nc = [g.reshape(-1) for g in np.meshgrid(lv, lv, lv, indexing='ij')]
    J_C, J_M = float(p['tm_J_C']), float(p['tm_J_M'])
    h = J_C * n2 + J_M * n3 - 0.5 * (J_C + J_M) * nc
    fwd, bwd = np.exp(0.5 * h), np.exp(-0.5 * h)
    k01 = p['tm_k_01'] * ca * fwd
    k12 = p['tm_k_12'] * fwd
    k23 = p['tm_k_23'] * fwd
    k30 = np.full(27, float(p['tm_k_30']))
    k10 = (p['tm_k_01'] / p['tm_Keq_01']) * bwd
    k21 = (p['tm_k_12'] / p['tm_Keq_12']) * bwd
    k32 = (p['tm_k_23'] / p['tm_Keq_23']) * bwd
    Q = np.zeros((27, 4, 4))
    Q[:, 0, 0], Q[:, 0, 1] = -k01, k01
    Q[:, 1, 0], Q[:, 1, 1], Q[:, 1, 2] = k10, -(k10 + k12), k12
    Q[:, 2, 1], Q[:, 2, 2], Q[:, 2, 3] = k21, -(k21 + k23), k23
    Q[:, 3, 0], Q[:, 3, 2], Q[:, 3, 3] = k30, k32, -(k30 + k32)
    return Q


def kT_pNnm(temp_celsius: float) -> float:
    """Thermal energy in pN*nm with k_B = 1.3810e-23 J/K (transitions.py:708)."""
    return 1.3810e-23 * (float(temp_celsius) + 273.15) * 1e21


def xb_generators(p: Mapping[str, float], pCa: float, d: float, x: np.ndarray,
                  perm: np.ndarray) -> np.ndarray:
    """6x6 XB generators at axial offsets x, radial d (transitions.py:632-804).

    Rate laws: rate_functions.py:318-753 (xb_rate_01 .. xb_rate_05).
    """
    x = np.asarray(x, dtype=np.float64)
    perm = np.broadcast_to(np.asarray(perm, dtype=np.float64), x.shape)
    d = float(d)
    r = np.sqrt(x * x + d * d)
    th = np.arctan2(d, x)
    kt = kT_pNnm(p['temp_celsius'])
    E_w = (0.5 * p['xb_g_k_weak'] * (r - p['xb_g_rest_weak']) ** 2
           + 0.5 * p['xb_c_k_weak'] * (th - p['xb_c_rest_weak']) ** 2) / kt
    E_s = (0.5 * p['xb_g_k_strong'] * (r - p['xb_g_rest_strong']) ** 2
           + 0.5 * p['xb_c_k_strong'] * (th - p['xb_c_rest_strong']) ** 2) / kt
    E_diff = E_w - E_s                                     # plain subtraction (:721)
    U_drx = float(p['xb_U_DRX'])
    U_loose = p['xb_U_loose'] + E_w
    U_t1 = p['xb_U_tight_1'] + E_s
    U_t2 = p['xb_U_tight_2'] + E_s
    f_strong = (p['xb_g_k_strong'] * (r - p['xb_g_rest_strong'])
                + (1.0 / r) * p['xb_c_k_strong'] * (th - p['xb_c_rest_strong']))   # (:737)
    ca = 10.0 ** (-float(pCa))
    with np.errstate(all='ignore'):
        r01 = perm * p['xb_r01_coeff'] * np.exp(-E_w)
        r01 = np.where(np.isnan(r01), 0.0, r01)
        r10 = np.minimum(np.exp(np.log(r01 + 0.005) - (U_drx - U_loose)), RATE_CAP)
        r10 = np.where(np.isnan(r10), RATE_CAP, r10)
        r12 = p['xb_r12_coeff'] * np.exp(np.minimum(E_diff, 30.0) / 2.0)
        r12 = np.where(np.isnan(r12), 0.0, r12)
        r21 = np.minimum(np.exp(np.log(r12) - (U_loose - U_t1)), RATE_CAP)
        r23 = np.minimum(p['xb_r23_coeff'] * np.exp(-f_strong * p['xb_delta_23'] / kt), RATE_CAP)
        r32 = np.exp(np.minimum(np.log(r23 + 1e-30) + (U_t2 - U_t1), np.log(RATE_CAP)))
        r34 = np.minimum(p['xb_r34_coeff'] * np.exp(f_strong * p['xb_delta_34'] / kt), RATE_CAP)
    r40 = float(p['xb_r40'])
    r04 = float(p['xb_r04'])
    b = float(p['xb_srx_b'])
    r50 = p['xb_srx_k0'] + ((p['xb_srx_kmax'] - p['xb_srx_k0']) * ca ** b) / (
        p['xb_srx_ca50'] ** b + ca ** b)
    r05 = float(p['xb_r05'])
    n = x.shape[0]
    Q = np.zeros((n, 6, 6))
    Q[:, 0, 0] = -(r01 + r04 + r05)
    Q[:, 0, 1], Q[:, 0, 4], Q[:, 0, 5] = r01, r04, r05
    Q[:, 1, 0], Q[:, 1, 1], Q[:, 1, 2] = r10, -(r10 + r12), r12
    Q[:, 2, 1], Q[:, 2, 2], Q[:, 2, 3] = r21, -(r21 + r23), r23
    Q[:, 3, 2], Q[:, 3, 3], Q[:, 3, 4] = r32, -(r32 + r34), r34
    Q[:, 4, 0], Q[:, 4, 3], Q[:, 4, 4] = r40, 0.0, -r40
    Q[:, 5, 0], Q[:, 5, 5] = r50, -r50
    return Q


def xb_bin_generators(p, pCa, d, bin_centers) -> np.ndarray:
    """(2*n_bins, 6, 6): permissiveness-0 block then -1 block (transitions.py:889-912)."""
    c = np.asarray(bin_centers, dtype=np.float64)
    return np.concatenate([xb_generators(p, pCa, d, c, 0.0),
                           xb_generators(p, pCa, d, c, 1.0)])


def absorbing(Q: np.ndarray) -> np.ndarray:
    """Rows 4 and 0 zeroed (transitions.py:1077)."""
    Q = Q.copy()
    Q[:, 4, :] = 0.0
    Q[:, 0, :] = 0.0
    return Q


@dataclass
class Tables:
    """Per-instance transition tables for one (pCa, d, dt)."""
    P_tm: np.ndarray     # (27, 4, 4)
    P_xb: np.ndarray     # (2*nb, 6, 6)
    P_abs: np.ndarray    # (2*nb, 6, 6)


@dataclass
class Pop:
    """Subpopulations of one instance (core/subpopulation.py).

    mode 'mean_field': one generator sum_k f_k Q_k (transitions.py:585-590,
    966-969); 'explicit': per-population tables stacked K deep, selected by
    the per-site / per-head labels (transitions.py:591-602, 971-975).
    params: K parameter mappings (0 = wild type); fractions: K weights
    (float32 values, as the device receives them); has_xb / has_tm: whether
    the XB / TM kinetics are population-dependent at all.
    """
    mode: str
    params: list
    fractions: np.ndarray
    has_xb: bool = True
    has_tm: bool = False
    xb_label: Optional[np.ndarray] = None   # (X,) explicit
    tm_label: Optional[np.ndarray] = None   # (S,) explicit

    def sets(self):
        ex = self.mode == 'explicit'
        return (len(self.params) if ex and self.has_tm else 1), (len(self.params) if ex and self.has_xb else 1)

    def offsets(self, nb: int):
        """Row offsets: TM config + 27 * label, XB key + 2 nb * label."""
        ktm, kxb = self.sets()
        tm = (27 * np.asarray(self.tm_label, np.int64)) if ktm > 1 else 0
        xb = (2 * nb * np.asarray(self.xb_label, np.int64)) if kxb > 1 else 0
        return tm, xb


def _pop_generators(gen, pop: Optional[Pop], flag: bool, p):
    """Generators of one kinetics: the wild type's, the mean-field blend, or
    the K population blocks stacked along the first axis."""
    if pop is None or not flag:
        return gen(p)
    Qs = [gen(pk) for pk in pop.params]
    if pop.mode == 'mean_field':
        f = np.asarray(pop.fractions, np.float32).astype(np.float64)
        return sum(f[k] * Qs[k] for k in range(len(Qs)))
    return np.concatenate(Qs)


def make_tables(p, pCa, d, dt, bin_centers, d_abs=None, pop: Optional[Pop] = None) -> Tables:
    """expm(Q dt) for TM, XB and the absorbing XB generator (transitions.py:
    597-598, 993-1026, 1029-1078); ``d_abs`` is the spacing the metrics path
    uses (simulation.py:711); ``pop`` the subpopulations."""
    has_tm = pop is not None and pop.has_tm
    has_xb = pop is not None and pop.has_xb
    P_tm = expm_ref(_pop_generators(lambda q: tm_generators(q, pCa), pop, has_tm, p) * dt)
    Q = _pop_generators(lambda q: xb_bin_generators(q, pCa, d, bin_centers), pop, has_xb, p)
    P_xb = expm_ref(Q * dt)
    if d_abs is not None and float(d_abs) != float(d):
        Q = _pop_generators(lambda q: xb_bin_generators(q, pCa, d_abs, bin_centers), pop, has_xb, p)
    P_abs = expm_ref(absorbing(Q) * dt)
    return Tables(P_tm, P_xb, P_abs)


# ============================================================================
# state
# ============================================================================

@dataclass
class State:
    """Per-instance state (state.py:69-113), flat.

    thick (Tk, nc) and thin (Tn, ns) positions are float64; xb_xdist is the
    float32 axial offset stored by the last nearest-site update (its radial
    column is the lattice spacing and is not stored).
    """
    thick: np.ndarray
    thin: np.ndarray
    xb_state: np.ndarray      # (X,) int8
    xb_bound: np.ndarray      # (X,) int32
    xb_nearest: np.ndarray    # (X,) int32
    xb_xdist: np.ndarray      # (X,) float32
    tm: np.ndarray            # (Tn, ns) int8
    thin_bound: np.ndarray    # (Tn, ns) int32
    ls: float = 14.0          # carried lattice spacing (dynamic mode)

    def copy(self) -> 'State':
        return State(self.thick.copy(), self.thin.copy(), self.xb_state.copy(),
                     self.xb_bound.copy(), self.xb_nearest.copy(), self.xb_xdist.copy(),
                     self.tm.copy(), self.thin_bound.copy(), float(self.ls))

    def x(self) -> np.ndarray:
        return np.concatenate([self.thick.ravel(), self.thin.ravel()])

    def set_x(self, x):
        nk = self.thick.size
        self.thick = np.asarray(x[:nk], np.float64).reshape(self.thick.shape).copy()
        self.thin = np.asarray(x[nk:], np.float64).reshape(self.thin.shape).copy()


def realize_state(topo, z0: float) -> State:
    """Initial state (state.py:191-262): crowns at offsets, thin at z0 - offsets."""
    X = topo.total_xbs
    thin32 = (np.float32(z0) - topo.binding_offsets.astype(np.float32)).astype(np.float32)
    return State(
        thick=topo.crown_offsets.astype(np.float32).astype(np.float64),
        thin=thin32.astype(np.float64),
        xb_state=np.zeros(X, np.int8), xb_bound=np.full(X, -1, np.int32),
        xb_nearest=np.zeros(X, np.int32), xb_xdist=np.zeros(X, np.float32),
        tm=np.zeros((topo.n_thin, topo.n_sites), np.int8),
        thin_bound=np.full((topo.n_thin, topo.n_sites), -1, np.int32))


# ============================================================================
# kinetics
# ============================================================================

def nearest_sites(thick32: np.ndarray, thin32: np.ndarray, topo) -> Tuple[np.ndarray, np.ndarray]:
    """Nearest reachable site per XB and its axial offset (geometry.py:60-262).

    Search from crown + 13 nm; mask padding slots and sites at x <= 0; argmin
    takes the first minimum; offset measured from the crown base.  float32.
    """
    nx = topo.n_xb_per_crown
    base = np.repeat(thick32.astype(np.float32), nx, axis=1).reshape(-1)
    head = (base + XB_REACH).astype(np.float32)
    tid = topo.xb_to_thin_id
    cand = topo.xb_to_site_indices
    cpos = thin32.astype(np.float32)[tid[:, None], cand]
    dist = np.abs(cpos - head[:, None]).astype(np.float32)
    npf = topo.n_sites_per_face[tid, topo.xb_to_thin_face]
    ok = (np.arange(cand.shape[1])[None, :] < npf[:, None]) & (cpos > np.float32(0.0))
    md = np.where(ok, dist, np.float32(np.inf))
    j = np.argmin(md, axis=1)
    nearest = cand[np.arange(cand.shape[0]), j].astype(np.int32)
    xdist = (thin32.astype(np.float32)[tid, nearest] - base).astype(np.float32)
    return nearest, xdist


def bin_index(xdist32: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """clip(digitize(x, edges) - 1, 0, nb-1) (transitions.py:915-917)."""
    nb = len(edges) - 1
    b = np.searchsorted(edges.astype(np.float32), xdist32.astype(np.float32), side='right') - 1
    return np.clip(b, 0, nb - 1)


def permissiveness(tm: np.ndarray, nearest: np.ndarray, topo) -> np.ndarray:
    """(tm[thin, clip(nearest)] == 3) * xb_valid (transitions.py:873-885)."""
    site = np.clip(nearest, 0, topo.n_sites - 1)
    return (tm[topo.xb_to_thin_id, site] == 3) & topo.xb_valid.astype(bool)


def xb_keys(xdist32, nearest, tm, topo) -> np.ndarray:
    """key = perm * n_bins + bin (transitions.py:919-922)."""
    nb = len(topo.xb_bin_edges) - 1
    return permissiveness(tm, nearest, topo).astype(np.int64) * nb + bin_index(xdist32, topo.xb_bin_edges)


def sample_rows(prob: np.ndarray, u: np.ndarray) -> np.ndarray:
    """argmax(u < cumsum(prob)); 0 when no entry is true (transitions.py:618-619)."""
    hit = u[:, None] < np.cumsum(prob, axis=1)
    return np.where(hit.any(axis=1), np.argmax(hit, axis=1), 0)


def tm_configs(tm: np.ndarray, topo) -> np.ndarray:
    """9*n2 + 3*n3 + n_closed over same-chain neighbours (transitions.py:384-431,575-578)."""
    prev, nxt = topo.tm_prev_neighbor, topo.tm_next_neighbor
    rows = np.arange(tm.shape[0])[:, None]
    own = np.arange(tm.shape[1])[None, :]
    tp, tn = tm[rows, prev], tm[rows, nxt]
    pr, nr = prev != own, nxt != own
    n2 = ((tp == 2) & pr).astype(int) + ((tn == 2) & nr).astype(int)
    n3 = ((tp == 3) & pr).astype(int) + ((tn == 3) & nr).astype(int)
    nc = ((tp <= 1) & pr).astype(int) + ((tn <= 1) & nr).astype(int)
    return (np.minimum(n2, 2) * 9 + np.minimum(n3, 2) * 3 + np.minimum(nc, 2)).reshape(-1)


def tm_transition(tm, thin_bound, P_tm, u_tm, topo, off=0) -> np.ndarray:
    """TM sampling with the bound-open lock (transitions.py:516-625); ``off``
    adds each site's population block (27 * label) to its row."""
    cfg = tm_configs(tm, topo) + off
    s = tm.reshape(-1).astype(np.int64)
    prob = P_tm[cfg, s].copy()
    locked = (s == 3) & (thin_bound.reshape(-1) >= 0)
    prob[locked] = np.array([0.0, 0.0, 0.0, 1.0])
    return sample_rows(prob, np.asarray(u_tm, np.float64)).astype(np.int8).reshape(tm.shape)


def xb_transition(st: State, tm_new, P_xb, u_xb, topo, off=0):
    """XB sampling and binding bookkeeping (transitions.py:1081-1245); ``off``
    adds each head's population block (2 n_bins * label) to its key.

    Returns (xb_state, xb_bound, thin_bound) after the step.
    """
    ns = topo.n_sites
    key = xb_keys(st.xb_xdist, st.xb_nearest, tm_new, topo) + off
    old = st.xb_state.astype(np.int64)
    new = sample_rows(P_xb[key, old], np.asarray(u_xb, np.float64))
    perm = permissiveness(tm_new, st.xb_nearest, topo)
    old_b = (old >= 1) & (old <= 3)
    new_b = (new >= 1) & (new <= 3)
    binding = ~old_b & new_b
    unbinding = old_b & ~new_b
    tid = topo.xb_to_thin_id
    near_site = np.clip(st.xb_nearest, 0, ns - 1)
    tb = st.thin_bound.reshape(-1)
    occupied = tb[tid * ns + near_site] >= 0
    can_bind = binding & perm & ~occupied
    xb_bound = np.where(can_bind, st.xb_nearest, np.where(unbinding, -1, st.xb_bound)).astype(np.int32)
    tb_new = tb.copy()
    clr = unbinding & (st.xb_bound >= 0)
    tb_new[tid[clr] * ns + np.clip(st.xb_bound[clr], 0, ns - 1)] = -1
    idx = np.nonzero(can_bind)[0]
    for i in idx[::-1]:            # descending, so the lowest index writes last
        tb_new[tid[i] * ns + near_site[i]] = i
    new = np.where(binding & ~can_bind, 0, new)
    return new.astype(np.int8), xb_bound, tb_new.reshape(st.thin_bound.shape).astype(np.int32)


# ============================================================================
# mechanics
# ============================================================================

def _xb_springs(p, strong):
    gk = np.where(strong, p['xb_g_k_strong'], p['xb_g_k_weak'])
    gr = np.where(strong, p['xb_g_rest_strong'], p['xb_g_rest_weak'])
    ck = np.where(strong, p['xb_c_k_strong'], p['xb_c_k_weak'])
    cr = np.where(strong, p['xb_c_rest_strong'], p['xb_c_rest_weak'])
    return gk, gr, ck, cr


@dataclass
class Mech:
    """Everything the residual needs besides the positions."""
    topo: object
    p: Mapping[str, float]
    z: float
    d: float
    xb_state: np.ndarray
    xb_bound: np.ndarray

    def __post_init__(self):
        t = self.topo
        self.nk = t.n_thick * t.n_crowns
        s = self.xb_state.astype(np.int64)
        self.bound = (s >= 1) & (s <= 3) & (self.xb_bound >= 0)
        self.strong = (s == 2) | (s == 3)
        self.crown = np.repeat(np.arange(self.nk), t.n_xb_per_crown)
        self.site = self.nk + (np.clip(t.xb_to_thin_id, 0, t.n_thin - 1) * t.n_sites
                               + np.clip(self.xb_bound, 0, t.n_sites - 1))


def xb_geometry(m: Mech, x: np.ndarray, d: float):
    xd = x[m.site] - x[m.crown]
    r = np.sqrt(xd * xd + d * d)
    rs = np.where(r > 1e-10, r, 1e-10)
    th = np.arctan2(d, xd)
    return xd, rs, th


def residual(m: Mech, x: np.ndarray, d: Optional[float] = None) -> np.ndarray:
    """Axial force residual F(x) (forces.py:82-500), float64."""
    t, p = m.topo, m.p
    d = m.d if d is None else d
    Tk, nc, Tn, ns = t.n_thick, t.n_crowns, t.n_thin, t.n_sites
    xk = x[:m.nk].reshape(Tk, nc)
    xn = x[m.nk:].reshape(Tn, ns)
    k, kn = float(p['thick_k']), float(p['thin_k'])
    # thick backbone + titin on the tip crown (forces.py:82-157)
    sk = (np.diff(np.concatenate([np.zeros((Tk, 1)), xk], axis=1), axis=1)
          - t.crown_rests.astype(np.float64)) * k
    ad = m.z - xk[:, -1]
    L = np.sqrt(ad * ad + d * d)
    F = np.maximum(p['titin_a'] * np.exp(np.clip(p['titin_b'] * (L - p['titin_rest']), -100, 100)), 0.0)
    cos = np.where(L > 0, ad / np.where(L > 0, L, 1.0), 0.0)
    tt = N_TITIN_PER_THICK * F * cos
    fk = np.diff(np.concatenate([sk, tt[:, None]], axis=1), axis=1)
    # thin backbone anchored at the Z-line (forces.py:210-270)
    sn = (np.diff(np.concatenate([xn, np.full((Tn, 1), m.z)], axis=1), axis=1)
          - t.binding_rests.astype(np.float64)) * kn
    fn = np.diff(np.concatenate([np.zeros((Tn, 1)), sn], axis=1), axis=1)
    out = np.concatenate([fk.ravel(), fn.ravel()])
    # crossbridges (forces.py:277-418)
    xd, r, th = xb_geometry(m, x, d)
    gk, gr, ck, cr = _xb_springs(p, m.strong)
    f = gk * (r - gr) * (xd / r) - (1.0 / r) * ck * (th - cr) * (d / r)
    f = np.where(m.bound, f, 0.0)
    np.add.at(out, m.crown, f)
    np.add.at(out, m.site, -f)
    return out


def xb_stiffness(m: Mech, x: np.ndarray, d: Optional[float] = None) -> np.ndarray:
    """df/dx_dist of each bound XB's axial force (0 for unbound)."""
    d = m.d if d is None else d
    xd, r, th = xb_geometry(m, x, d)
    gk, gr, ck, cr = _xb_springs(m.p, m.strong)
    kx = gk * (1.0 - gr * d * d / r ** 3) + ck * d * (d + 2.0 * xd * (th - cr)) / r ** 4
    return np.where(m.bound, kx, 0.0)


def titin_tip_stiffness(m: Mech, x: np.ndarray, d: Optional[float] = None) -> np.ndarray:
    """d(titin axial force on the tip)/d(tip position), per thick filament."""
    t, p = m.topo, m.p
    d = m.d if d is None else d
    xk = x[:m.nk].reshape(t.n_thick, t.n_crowns)
    ad = m.z - xk[:, -1]
    L = np.sqrt(ad * ad + d * d)
    arg = p['titin_b'] * (L - p['titin_rest'])
    F = p['titin_a'] * np.exp(np.clip(arg, -100, 100))
    inrange = (arg > -100) & (arg < 100)
    Fp = np.where(inrange, p['titin_b'] * F, 0.0)
    return -N_TITIN_PER_THICK * (Fp * ad * ad / L ** 2 + F * d * d / L ** 3)


def jvp(m: Mech, kx: np.ndarray, ktit: np.ndarray, v: np.ndarray) -> np.ndarray:
    """J v for the axial residual (what jax.jvp yields at solver.py:413)."""
    t, p = m.topo, m.p
    Tk, nc, Tn, ns = t.n_thick, t.n_crowns, t.n_thin, t.n_sites
    k, kn = float(p['thick_k']), float(p['thin_k'])
    vk = v[:m.nk].reshape(Tk, nc)
    vn = v[m.nk:].reshape(Tn, ns)
    dsk = k * np.diff(np.concatenate([np.zeros((Tk, 1)), vk], axis=1), axis=1)
    fk = np.diff(np.concatenate([dsk, (ktit * vk[:, -1])[:, None]], axis=1), axis=1)
    dsn = kn * np.diff(np.concatenate([vn, np.zeros((Tn, 1))], axis=1), axis=1)
    fn = np.diff(np.concatenate([np.zeros((Tn, 1)), dsn], axis=1), axis=1)
    out = np.concatenate([fk.ravel(), fn.ravel()])
    q = kx * (v[m.site] - v[m.crown])
    np.add.at(out, m.crown, q)
    np.add.at(out, m.site, -q)
    return out


def jacobian_sparse(m: Mech, x: np.ndarray, d: Optional[float] = None):
    """Assembled axial Jacobian (scipy CSR), same derivatives as ``jvp``."""
    import scipy.sparse as sps
    t, p = m.topo, m.p
    Tk, nc, Tn, ns = t.n_thick, t.n_crowns, t.n_thin, t.n_sites
    N = m.nk + Tn * ns
    k, kn = float(p['thick_k']), float(p['thin_k'])
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(np.atleast_1d(r)); cols.append(np.atleast_1d(c)); vals.append(np.atleast_1d(v))

    kx = xb_stiffness(m, x, d)
    ktit = titin_tip_stiffness(m, x, d)
    for f in range(Tk):
        o = f * nc
        i = np.arange(nc)
        diag = np.full(nc, -2.0 * k)
        diag[-1] = -k
        add(o + i, o + i, diag)
        add(o + i[:-1], o + i[1:], np.full(nc - 1, k))
        add(o + i[1:], o + i[:-1], np.full(nc - 1, k))
        add(o + nc - 1, o + nc - 1, ktit[f])
    for f in range(Tn):
        o = m.nk + f * ns
        i = np.arange(ns)
        diag = np.full(ns, -2.0 * kn)
        diag[0] = -kn
        add(o + i, o + i, diag)
        add(o + i[:-1], o + i[1:], np.full(ns - 1, kn))
        add(o + i[1:], o + i[:-1], np.full(ns - 1, kn))
    c, s = m.crown, m.site
    add(c, s, kx); add(c, c, -kx); add(s, s, -kx); add(s, c, kx)
    R = np.concatenate(rows); C = np.concatenate(cols); V = np.concatenate(vals)
    return sps.csr_matrix((V, (R, C)), shape=(N, N))


# ---------------------------------------------------------------------------
# preconditioner
# ---------------------------------------------------------------------------

def thomas_factor(lower, diag, upper):
    """Forward sweep (solver.py:128-162)."""
    n = len(diag)
    md = np.empty(n)
    mult = np.empty(n - 1)
    md[0] = diag[0]
    for i in range(1, n):
        mult[i - 1] = lower[i - 1] / md[i - 1]
        md[i] = diag[i] - mult[i - 1] * upper[i - 1]
    return 1.0 / md, np.asarray(upper, np.float64), mult


def thomas_solve(fac, rhs):
    """Forward + back substitution, sequential, over the last axis (solver.py:165-219)."""
    inv, up, mult = fac
    y = np.array(rhs, dtype=np.float64)
    n = y.shape[-1]
    for i in range(1, n):
        y[..., i] -= mult[i - 1] * y[..., i - 1]
    xx = np.empty_like(y)
    xx[..., n - 1] = y[..., n - 1] * inv[n - 1]
    for i in range(n - 2, -1, -1):
        xx[..., i] = (y[..., i] - up[i] * xx[..., i + 1]) * inv[i]
    return xx


def preconditioner(topo, p, eps=1e-9):
    """Negated backbone tridiagonals + eps, factored once (state.py:269-316;
    solver.py:243-282).  Returns v -> M^-1 v for the axial DOF vector."""
    nc, ns = topo.n_crowns, topo.n_sites
    k, kn = float(p['thick_k']), float(p['thin_k'])
    dk = np.full(nc, 2.0 * k); dk[-1] = k
    dn = np.full(ns, 2.0 * kn); dn[0] = kn
    fk = thomas_factor(np.full(nc - 1, -k), dk + eps, np.full(nc - 1, -k))
    fn = thomas_factor(np.full(ns - 1, -kn), dn + eps, np.full(ns - 1, -kn))
    nk = topo.n_thick * nc

    def apply(v):
        a = thomas_solve(fk, v[:nk].reshape(topo.n_thick, nc)).ravel()
        b = thomas_solve(fn, v[nk:].reshape(topo.n_thin, ns)).ravel()
        return np.concatenate([a, b])
    return apply


def pcg(amul, pmul, b, n_cg):
    """Unrolled preconditioned CG from x0 = 0 (solver.py:329-366)."""
    x = np.zeros_like(b)
    r = b.copy()             # b - A x0 with x0 = 0
    z = pmul(r)
    pv = z.copy()
    rz = float(r @ z)
    for _ in range(n_cg):
        Ap = amul(pv)
        alpha = rz / (float(pv @ Ap) + 1e-30)
        x = x + alpha * pv
        r = r - alpha * Ap
        z = pmul(r)
        rz_new = float(r @ z)
        beta = rz_new / (rz + 1e-30)
        pv = z + beta * pv
        rz = rz_new
    return x


def solver_tolerance(p, tol=None) -> float:
    """max(tol or solver_tol, thick_k * 1e-4, 0.25) (solver.py:714-718), float32."""
    t = np.float32(p['solver_tol'] if tol is None else tol)
    floor = np.float32(np.float32(p['thick_k']) * np.float32(1e-4))
    return float(max(t, max(floor, np.float32(MIN_FLOAT32_TOLERANCE))))


def newton_cg(m: Mech, x0, precond, tol, n_newton, n_cg):
    """Newton with unrolled PCG on -J (solver.py:369-427).  Returns (x, iters, max|F|)."""
    x = np.array(x0, dtype=np.float64)
    f = residual(m, x)
    it = 0
    while np.max(np.abs(f)) > tol and it < n_newton:
        kx = xb_stiffness(m, x)
        kt = titin_tip_stiffness(m, x)
        dx = pcg(lambda v: -jvp(m, kx, kt, v), precond, f, n_cg)
        dx = np.where(np.isfinite(dx), dx, 0.0)
        x = x + dx
        f = residual(m, x)
        it += 1
    return x, it, float(np.max(np.abs(f)))


def newton_exact(m: Mech, x0, tol=1e-8, max_iter=50):
    """Newton with a direct sparse solve, to a tight tolerance (independent of CG)."""
    import scipy.sparse.linalg as spla
    x = np.array(x0, dtype=np.float64)
    f = residual(m, x)
    it = 0
    while np.max(np.abs(f)) > tol and it < max_iter:
        J = jacobian_sparse(m, x)
        x = x - spla.spsolve(J.tocsc(), f)
        f = residual(m, x)
        it += 1
    return x, it, float(np.max(np.abs(f)))


# ---------------------------------------------------------------------------
# dynamic lattice spacing (solver.py:502-656; forces.py:649-749)
# ---------------------------------------------------------------------------

def radial_residual(m: Mech, x, d, K_lat, d_ref) -> float:
    """F_rad = -K_lat (d - d_ref) - sum f_xb,rad - sum f_titin,rad (solver.py:502-528)."""
    t, p = m.topo, m.p
    xd, r, th = xb_geometry(m, x, d)
    gk, gr, ck, cr = _xb_springs(p, m.strong)
    g = gk * (r - gr) * (d / r) + (1.0 / r) * ck * (th - cr) * (xd / r)
    g = np.where(m.bound, g, 0.0).sum()
    xk = x[:m.nk].reshape(t.n_thick, t.n_crowns)
    ad = m.z - xk[:, -1]
    L = np.sqrt(ad * ad + d * d)
    Ls = np.where(L > 1e-10, L, 1e-10)
    F = np.maximum(p['titin_a'] * np.exp(np.clip(p['titin_b'] * (L - p['titin_rest']), -100, 100)), 0.0)
    tr = (N_TITIN_PER_THICK * F * (d / Ls)).sum()
    return float(-K_lat * (d - d_ref) - g - tr)


def aug_residual(m: Mech, xa, K_lat, d_ref):
    d = float(xa[-1])
    return np.concatenate([residual(m, xa[:-1], d), [radial_residual(m, xa[:-1], d, K_lat, d_ref)]])


def aug_partials(m: Mech, x, d):
    """Partial derivatives for the bordered Jacobian at (x, d).

    Returns (dFax_dd (N,), dFrad_dx (N,), dFrad_dd_without_Klat).
    """
    t, p = m.topo, m.p
    N = x.shape[0]
    xd, r, th = xb_geometry(m, x, d)
    gk, gr, ck, cr = _xb_springs(p, m.strong)
    r3, r4 = r ** 3, r ** 4
    f_d = gk * gr * xd * d / r3 - ck * (xd * d + (th - cr) * (xd * xd - d * d)) / r4
    g_x = f_d
    g_d = gk * (1.0 - gr * xd * xd / r3) + ck * xd * (xd - 2.0 * d * (th - cr)) / r4
    f_d = np.where(m.bound, f_d, 0.0)
    g_x = np.where(m.bound, g_x, 0.0)
    g_d = np.where(m.bound, g_d, 0.0)
    dFax_dd = np.zeros(N)
    np.add.at(dFax_dd, m.crown, f_d)
    np.add.at(dFax_dd, m.site, -f_d)
    dFrad_dx = np.zeros(N)
    np.add.at(dFrad_dx, m.crown, g_x)     # -d(sum g)/dx_crown, x_dist = site - crown
    np.add.at(dFrad_dx, m.site, -g_x)
    xk = x[:m.nk].reshape(t.n_thick, t.n_crowns)
    ad = m.z - xk[:, -1]
    L = np.sqrt(ad * ad + d * d)
    arg = p['titin_b'] * (L - p['titin_rest'])
    F = p['titin_a'] * np.exp(np.clip(arg, -100, 100))
    bF = np.where((arg > -100) & (arg < 100), p['titin_b'] * F, 0.0)
    tt_d = N_TITIN_PER_THICK * ad * d * (bF / L ** 2 - F / L ** 3)
    tr_d = N_TITIN_PER_THICK * (bF * d * d / L ** 2 + F * ad * ad / L ** 3)
    tr_p = -tt_d
    tips = np.arange(t.n_thick) * t.n_crowns + t.n_crowns - 1
    dFax_dd[tips] += tt_d
    dFrad_dx[tips] += -tr_p
    return dFax_dd, dFrad_dx, float(-g_d.sum() - tr_d.sum())


def newton_cg_dynamic(m: Mech, x0, d0, K_lat, d_ref, precond, tol, n_newton, n_cg):
    """Bordered (N+1) Newton-CG with the d >= 1 nm clamp (solver.py:599-656)."""
    _, _, jdd_rest = aug_partials(m, x0, d0)
    J_dd = -K_lat + jdd_rest
    d_inv = -1.0 / J_dd
    xa = np.concatenate([np.asarray(x0, np.float64), [float(d0)]])
    f = aug_residual(m, xa, K_lat, d_ref)
    it = 0

    def pmul(v):
        return np.concatenate([precond(v[:-1]), [d_inv * v[-1]]])

    while np.max(np.abs(f)) > tol and it < n_newton:
        x, d = xa[:-1], float(xa[-1])
        kx = xb_stiffness(m, x, d)
        kt = titin_tip_stiffness(m, x, d)
        fax_d, frad_x, frad_d = aug_partials(m, x, d)
        frad_d = frad_d - K_lat

        def amul(v):
            top = jvp(m, kx, kt, v[:-1]) + fax_d * v[-1]
            bot = float(frad_x @ v[:-1]) + frad_d * v[-1]
            return -np.concatenate([top, [bot]])

        dx = pcg(amul, pmul, f, n_cg)
        dx = np.where(np.isfinite(dx), dx, 0.0)
        xa = xa + dx
        xa[-1] = max(xa[-1], 1.0)
        f = aug_residual(m, xa, K_lat, d_ref)
        it += 1
    return xa[:-1], float(xa[-1]), it, float(np.max(np.abs(f)))


def newton_exact_dynamic(m: Mech, x0, d0, K_lat, d_ref, tol=1e-8, max_iter=50):
    """Bordered Newton with a direct sparse solve (tight reference)."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    xa = np.concatenate([np.asarray(x0, np.float64), [float(d0)]])
    f = aug_residual(m, xa, K_lat, d_ref)
    it = 0
    while np.max(np.abs(f)) > tol and it < max_iter:
        x, d = xa[:-1], float(xa[-1])
        J = jacobian_sparse(m, x, d)
        fax_d, frad_x, frad_d = aug_partials(m, x, d)
        Ja = sps.bmat([[J, sps.csr_matrix(fax_d[:, None])],
                       [sps.csr_matrix(frad_x[None, :]), sps.csr_matrix([[frad_d - K_lat]])]])
        xa = xa - spla.spsolve(Ja.tocsc(), f)
        xa[-1] = max(xa[-1], 1.0)
        f = aug_residual(m, xa, K_lat, d_ref)
        it += 1
    return xa[:-1], float(xa[-1]), it, float(np.max(np.abs(f)))


def axial_force(thick: np.ndarray, topo, thick_k: float) -> float:
    """sum_t (x[t,0] - crown_offsets[t,0]) * thick_k (forces.py:588-627), float64."""
    return float(np.sum((thick[:, 0] - topo.crown_offsets[:, 0].astype(np.float64)) * float(thick_k)))


# ============================================================================
# metrics (metrics_fn.py:154-417)
# ============================================================================

def metrics(old: State, new: State, topo, p, z, pCa, ls, pre_thick, force, resid, iters,
            P_abs, key_old) -> Dict[str, float]:
    ox, nx = old.xb_state.astype(np.int64), new.xb_state.astype(np.int64)
    ntm = new.tm.astype(np.int64)
    X, S = float(nx.size), float(ntm.size)
    m = {}
    m['axial_force'] = force
    m['solver_residual'] = resid
    m['z_line'], m['pCa'], m['lattice_spacing'] = float(z), float(pCa), float(ls)
    cnt = [float(np.sum(nx == s)) for s in range(6)]
    m['n_bound'] = float(np.sum((nx >= 1) & (nx <= 3)))
    for nm, c in zip(('drx', 'loose', 'tight_1', 'tight_2', 'free_2', 'srx'), cnt):
        m['n_xb_' + nm] = c
    m['frac_xb_bound'] = m['n_bound'] / X
    for nm, c in zip(('drx', 'loose', 'tight_1', 'tight_2', 'free_2', 'srx'), cnt):
        m['frac_xb_' + nm] = c / X
    tcnt = [float(np.sum(ntm == s)) for s in range(4)]
    for s in range(4):
        m[f'n_tm_state_{s}'] = tcnt[s]
    for s in range(4):
        m[f'frac_tm_state_{s}'] = tcnt[s] / S
    m['actin_permissiveness'] = tcnt[3] / S
    co = topo.crown_offsets.astype(np.float32)
    near_b = float(co.min()) - 13.0
    far_b = float(co.max()) + 13.0
    ov = (new.thin >= near_b) & (new.thin <= far_b) & (new.thin > 0.0)
    nov = float(ov.sum())
    n2 = float(((ntm == 2) & ov).sum())
    n3 = float(((ntm == 3) & ov).sum())
    with np.errstate(divide='ignore', invalid='ignore'):
        m['frac_tm_state_2_overlap'] = n2 / nov if nov else float('nan')
        m['frac_tm_state_3_overlap'] = n3 / nov if nov else float('nan')
        m['frac_tm_available_overlap'] = (n2 + n3) / nov if nov else float('nan')
    m['n_overlap_sites'] = nov
    m['atp_consumed'] = float(np.sum((ox == 3) & ((nx == 4) | (nx == 0))))
    m['newly_bound'] = float(np.sum((ox == 0) & (nx == 1)))
    tdisp = (new.thick - co.astype(np.float64)).ravel()
    ndisp = (new.thin - np.cumsum(topo.binding_rests.astype(np.float32), axis=1).astype(np.float64)).ravel()
    for nm, a in (('thick', tdisp), ('thin', ndisp)):
        m[nm + '_displace_mean'] = float(a.mean())
        m[nm + '_displace_max'] = float(a.max())
        m[nm + '_displace_min'] = float(a.min())
        m[nm + '_displace_std'] = float(a.std())
    k = float(p['thick_k'])
    L0 = co[:, 0].astype(np.float64)
    e_new = 0.5 * k * (new.thick[:, 0] - L0) ** 2
    e_old = 0.5 * k * (old.thick[:, 0] - L0) ** 2
    m['thick_energy_first_avg'] = float(e_new.mean())
    m['thick_energy_first_delta_avg'] = float((e_new - e_old).mean())
    a, b, L0t = float(p['titin_a']), float(p['titin_b']), float(p['titin_rest'])

    def tit_energy(tip):
        L = np.sqrt((z - tip) ** 2 + ls ** 2)
        return (a / b) * (np.exp(b * (L - L0t)) - 1.0)
    tn, to = tit_energy(new.thick[:, -1]), tit_energy(old.thick[:, -1])
    m['titin_energy_avg'] = float(tn.mean())
    m['titin_energy_delta_avg'] = float((tn - to).mean())
    dx = new.thick - pre_thick
    work = force * float(dx.mean())
    m['work_thick'] = work
    m['work_thick_mean'] = work / float(topo.n_thick * topo.n_crowns)
    Pk = P_abs[key_old]
    atp = float(np.sum(np.where(ox == 3, Pk[:, 3, 4], 0.0)))
    tear = float(np.sum(np.where((ox == 2) | (ox == 3), Pk[np.arange(len(ox)), ox, 0], 0.0)))
    m['atp_expected_p'] = atp
    m['xb_tear_expected'] = tear
    m['work_per_atp'] = work / atp if atp > 0.01 else 0.0
    m['newton_iters'] = float(iters)
    return {kk: m[kk] for kk in METRIC_NAMES}


# ============================================================================
# the step and the run (timestep.py:65-199; simulation.py:633-747)
# ============================================================================

@dataclass
class SolverConfig:
    n_newton: int = 4
    n_cg: int = 6
    mode: str = "reference"        # "reference" (caps + tol floor) or "exact"
    exact_tol: float = 1e-8


def solve(st: State, topo, p, z, d, cfg: SolverConfig, precond,
          K_lat=None, d_ref=None, tol=None):
    """solve_equilibrium (solver.py:663-771).  Returns (x, d, iters, resid)."""
    m = Mech(topo, p, float(z), float(d), st.xb_state, st.xb_bound)
    x0 = st.x()
    if K_lat is None:
        if cfg.mode == "exact":
            x, it, res = newton_exact(m, x0, cfg.exact_tol)
        else:
            x, it, res = newton_cg(m, x0, precond, solver_tolerance(p, tol), cfg.n_newton, cfg.n_cg)
        return x, float(d), it, res
    if cfg.mode == "exact":
        x, dd, it, res = newton_exact_dynamic(m, x0, d, K_lat, d_ref, cfg.exact_tol)
    else:
        x, dd, it, res = newton_cg_dynamic(m, x0, d, K_lat, d_ref, precond,
                                           solver_tolerance(p, tol), cfg.n_newton, cfg.n_cg)
    return x, dd, it, res


def kinetics(st: State, topo, tables: Tables, u_tm, u_xb, pop: Optional[Pop] = None) -> State:
    """kinetics_step (timestep.py:65-129): nearest sites, TM, XB."""
    s = st.copy()
    off_tm, off_xb = pop.offsets(topo.n_xb_bins) if pop is not None else (0, 0)
    s.xb_nearest, s.xb_xdist = nearest_sites(s.thick.astype(np.float32),
                                             s.thin.astype(np.float32), topo)
    tm_new = tm_transition(s.tm, s.thin_bound, tables.P_tm, u_tm, topo, off_tm)
    s.xb_state, s.xb_bound, s.thin_bound = xb_transition(s, tm_new, tables.P_xb, u_xb, topo, off_xb)
    s.tm = tm_new
    return s


def init_state(topo, p, z0, cfg: SolverConfig, precond=None) -> State:
    """create_and_equilibrate (simulation.py:633-642).

    The solve uses the parameter row's z_line / lattice_spacing, not z0: the
    reference's run() never writes its z_line into the parameters, so for
    z0 != z_line the start is equilibrated against the parameter Z-line.
    """
    st = realize_state(topo, z0)
    st.xb_nearest, st.xb_xdist = nearest_sites(st.thick.astype(np.float32),
                                               st.thin.astype(np.float32), topo)
    pre = precond if precond is not None else preconditioner(topo, p)
    x, _, _, _ = solve(st, topo, p, float(p['z_line']), float(p['lattice_spacing']), cfg, pre)
    st.set_x(x)
    st.ls = float(np.float32(p['lattice_spacing']))
    return st


def step(st: State, topo, p, tables: Tables, z, pCa, ls_drv, dz, u_tm, u_xb, cfg: SolverConfig,
         precond, dynamic=None, pop: Optional[Pop] = None):
    """One scan body (simulation.py:679-721).

    dynamic: None for fixed spacing, else (K_lat, d_ref).  Returns
    (new_state, metrics dict).
    """
    old = st.copy()
    old.thin = old.thin + float(np.float32(dz))
    pre_thick = old.thick.copy()
    key_old = xb_keys(old.xb_xdist, old.xb_nearest, old.tm, topo) + \
        (pop.offsets(topo.n_xb_bins)[1] if pop is not None else 0)
    d_kin = old.ls if dynamic is not None else float(ls_drv)
    s = kinetics(old, topo, tables, u_tm, u_xb, pop)
    if dynamic is None:
        x, d_new, it, res = solve(s, topo, p, z, d_kin, cfg, precond)
    else:
        x, d_new, it, res = solve(s, topo, p, z, d_kin, cfg, precond,
                                  K_lat=dynamic[0], d_ref=dynamic[1])
    s.set_x(x)
    s.ls = d_new
    force = axial_force(s.thick, topo, p['thick_k'])
    mt = metrics(old, s, topo, p, z, pCa, d_new, pre_thick, force, res, it,
                 tables.P_abs, key_old)
    return s, mt


def run(topo, p, z_trace, pCa_trace, ls_trace, dt, seed, cfg: SolverConfig,
        K_lat=None, nu=0.0, n_steps=None, tables_fn=None, pop: Optional[Pop] = None):
    """run_single_sim for one instance (simulation.py:644-729), with the
    device's Philox uniforms.  Returns (final_state, metrics (T, 52))."""
    from .philox import uniforms, PURPOSE_TM, PURPOSE_XB
    z_trace = np.asarray(z_trace, np.float32)
    pCa_trace = np.asarray(pCa_trace, np.float32)
    ls_trace = np.asarray(ls_trace, np.float32)
    T = len(z_trace) if n_steps is None else n_steps
    pre = preconditioner(topo, p)
    st = init_state(topo, p, float(z_trace[0]), cfg, pre)
    st.ls = float(ls_trace[0])
    dz = np.concatenate([[np.float32(0)], np.diff(z_trace).astype(np.float32)])
    S, X = topo.n_thin * topo.n_sites, topo.total_xbs
    out = np.zeros((T, len(METRIC_NAMES)))
    cache = {}
    for t in range(T):
        dyn = None
        d_tab = float(ls_trace[t])
        if K_lat is not None:
            d_ref = float(ls_trace[t]) * (float(z_trace[0]) / float(z_trace[t])) ** float(nu)
            dyn = (float(K_lat), d_ref)
            d_tab = st.ls
        key = (float(pCa_trace[t]), d_tab)
        if K_lat is None and key in cache:
            tab = cache[key]
        else:
            tab = make_tables(p, pCa_trace[t], d_tab, dt, topo.xb_bin_centers, pop=pop)
            cache = {key: tab}
        if K_lat is not None:
            tab_m = None   # metrics tables use the solved spacing; patched below
        u_tm = uniforms(seed, t, PURPOSE_TM, S)
        u_xb = uniforms(seed, t, PURPOSE_XB, X)
        if K_lat is None:
            st, mt = step(st, topo, p, tab, z_trace[t], pCa_trace[t], ls_trace[t], dz[t],
                          u_tm, u_xb, cfg, pre, pop=pop)
        else:
            st_prev = st
            st, mt = step(st, topo, p, tab, z_trace[t], pCa_trace[t], ls_trace[t], dz[t],
                          u_tm, u_xb, cfg, pre, dynamic=dyn, pop=pop)
            tab_m = make_tables(p, pCa_trace[t], st.ls, dt, topo.xb_bin_centers, pop=pop)
            old = st_prev.copy()
            key_old = xb_keys(old.xb_xdist, old.xb_nearest, old.tm, topo) + \
                (pop.offsets(topo.n_xb_bins)[1] if pop is not None else 0)
            Pk = tab_m.P_abs[key_old]
            ox = old.xb_state.astype(np.int64)
            atp = float(np.sum(np.where(ox == 3, Pk[:, 3, 4], 0.0)))
            tear = float(np.sum(np.where((ox == 2) | (ox == 3), Pk[np.arange(len(ox)), ox, 0], 0.0)))
            mt['atp_expected_p'] = atp
            mt['xb_tear_expected'] = tear
            mt['work_per_atp'] = mt['work_thick'] / atp if atp > 0.01 else 0.0
        out[t] = [mt[k] for k in METRIC_NAMES]
    return st, out
