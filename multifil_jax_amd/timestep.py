"""The reference's single-step API on the device: ``timestep``,
``kinetics_step`` and ``solve_equilibrium`` (timestep.py:65-199,
kernels/solver.py:663-771), public in the reference "for future coupling"
(timestep.py:13-18), plus ``realize_state`` / ``create_and_equilibrate``
(core/state.py:191-262, simulation.py:633-642).

Each call is one launch of the engine's step kernel with ``n_steps = 1``
over a one-instance batch: the whole step, its kinetics half alone
(MF_LAUNCH_KINETICS_ONLY) or its solve alone (MF_LAUNCH_SOLVE_ONLY).
States travel as the reference's NamedTuples of numpy arrays.

Differences from the reference, each deliberate:

* ``rng_key`` is a uint64 Philox4x32-10 key instead of a threefry key: a
  step draws its uniforms from Philox under that key, and the returned key
  is the SplitMix64 successor (``split_key``).  ``uniforms=(u_tm, u_xb)``
  injects the draws instead (the reference's ``random_values=`` seam,
  transitions.py:521,1086).
* Axial positions come back as float64: the refined solver carries them in
  double, and a float32 round trip between calls would add ~1e-4 nm of
  rounding (~0.5 pN at thick_k 7500) to every step.
* ``solver`` selects the engine's solver mode ("reference": the FP32
  Newton-CG with the reference's caps and tolerance floor; "refined": FP64
  residual to 1e-5 pN).
* ``precond_params`` / ``prefactored_precond`` are accepted and unused: the
  engine applies the backbone preconditioner in closed form.
"""
from __future__ import annotations

from typing import NamedTuple, Optional, Tuple

import numpy as np

from .params import DYNAMIC_FIELDS, DynamicParams


class ThickState(NamedTuple):
    """Thick filament arrays (core/state.py:69-82)."""
    axial: np.ndarray           # (n_thick, n_crowns) crown positions (float64 here)
    xb_states: np.ndarray       # (n_thick, n_crowns, n_xb_per_crown) int8, 0-5
    xb_bound_to: np.ndarray     # same shape, int32 site index, -1 unbound
    xb_nearest_bs: np.ndarray   # same shape, int32 nearest site of the last update
    xb_distances: np.ndarray    # (..., 2) float32 (axial offset, radial = spacing)


class ThinState(NamedTuple):
    """Thin filament arrays (core/state.py:85-95)."""
    axial: np.ndarray           # (n_thin, n_sites) site positions (float64 here)
    tm_states: np.ndarray       # (n_thin, n_sites) int8, 0-3
    bound_to: np.ndarray        # (n_thin, n_sites) int32 bound XB, -1 none


class State(NamedTuple):
    """Pure simulation state (core/state.py:98-114)."""
    thick: ThickState
    thin: ThinState


class Drivers(NamedTuple):
    """Per-step overrides; NaN means "use the constant" (core/state.py:116-124)."""
    pCa: float
    z_line: float
    lattice_spacing: float


def resolve_value(driver_val, constant_val):
    """Driver if not NaN, else the constant (core/state.py:331-343)."""
    d = np.asarray(driver_val, np.float32)
    return np.where(np.isnan(d), np.asarray(constant_val, np.float32), d)


def split_key(key) -> np.uint64:
    """Successor of a Philox key (SplitMix64 finaliser of key + golden gamma)."""
    z = (int(key) + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return np.uint64(z ^ (z >> 31))


def realize_state(topology, z_line: float) -> State:
    """Initial state (core/state.py:191-262): crowns at their offsets, thin
    sites at z_line - binding offsets (float32), every head and site in state
    0 and unbound."""
    t = topology
    shp = (t.n_thick, t.n_crowns, t.n_xb_per_crown)
    thin = (np.float32(z_line) - np.asarray(t.binding_offsets, np.float32)).astype(np.float64)
    return State(
        ThickState(np.asarray(t.crown_offsets, np.float32).astype(np.float64),
                   np.zeros(shp, np.int8), np.full(shp, -1, np.int32),
                   np.zeros(shp, np.int32), np.zeros(shp + (2,), np.float32)),
        ThinState(thin.reshape(t.n_thin, t.n_sites), np.zeros((t.n_thin, t.n_sites), np.int8),
                  np.full((t.n_thin, t.n_sites), -1, np.int32)))


# ---------------------------------------------------------------------------
# one-instance engine context, cached per topology and solver shape
# ---------------------------------------------------------------------------
_ctx = {}


def _pack(topology):
    from .engine import LatticePack
    key = ("pack", id(topology))
    hit = _ctx.get(key)
    if hit is None or hit[0] is not topology:
        hit = (topology, LatticePack([topology]))
        _ctx[key] = hit
    return hit[1]


def _row(constants) -> np.ndarray:
    if isinstance(constants, DynamicParams):
        return constants.rows(1)
    return np.array([[float(constants[n]) for n in DYNAMIC_FIELDS]], np.float32)


def _populations(constants_row, xb_subpop, tm_subpop, topology):
    """The reference's (mode, constants_k, extra) tuples (simulation.py:
    659-675) as the engine's Populations of one instance."""
    if xb_subpop is None and tm_subpop is None:
        return None
    from .engine import Populations
    mode, ck, _ = xb_subpop if xb_subpop is not None else tm_subpop
    for sp in (xb_subpop, tm_subpop):
        if sp is not None and (sp[0] != mode or len(sp[1]) != len(ck)):
            raise ValueError("xb_subpop and tm_subpop must share the mode and the population constants")
    P = np.concatenate([_row(c) for c in ck])[None]
    P[0, 0] = constants_row[0]
    X, S = topology.total_xbs, topology.n_thin * topology.n_sites
    if mode == "mean_field":
        frac = np.asarray((xb_subpop or tm_subpop)[2], np.float32)[None]
        return Populations("mean_field", P, frac, has_xb=xb_subpop is not None, has_tm=tm_subpop is not None)
    if mode != "explicit":
        raise ValueError(f"unknown subpopulation mode {mode!r}")
    xl = np.asarray(xb_subpop[2], np.uint8).reshape(-1) if xb_subpop is not None else np.zeros(X, np.uint8)
    tl = np.asarray(tm_subpop[2], np.uint8).reshape(-1) if tm_subpop is not None else np.zeros(S, np.uint8)
    return Populations("explicit", P, xb_labels=[xl], tm_labels=[tl],
                       has_xb=xb_subpop is not None, has_tm=tm_subpop is not None)


def _batch(topology, row, drv, dynamic, pops):
    """A one-instance, one-step Batch; reused while the lattice, the spacing
    mode and the populations stay the same (parameters and drivers are
    updated in place)."""
    from .engine import Batch
    pack = _pack(topology)
    key = ("batch", id(topology), dynamic is not None,
           None if dynamic is None else float(dynamic))
    hit = _ctx.get(key)
    if pops is None and hit is not None and hit[0] is topology:
        bt = hit[1]
        bt.update_inputs(params=row, drivers=drv)
        return bt
    bt = Batch(pack, np.zeros(1, np.int32), row, drv, seeds=np.zeros(1, np.uint64),
               dynamic=None if dynamic is None else (np.float32(dynamic), np.float32(0.0)), pops=pops)
    if pops is None:
        _ctx[key] = (topology, bt)
    return bt


def _to_record(state: State, topology, ls: float, key_tm=None) -> dict:
    """A State as the engine's state record.  The record's table key of the
    previous step (used only by that step's metrics) is the bin of the
    stored axial distance at the stored nearest site, permissiveness from
    the current TM states (transitions.py:873-922)."""
    t = topology
    th, tn = state.thick, state.thin
    near = np.asarray(th.xb_nearest_bs, np.int64).reshape(-1)
    xd = np.asarray(th.xb_distances, np.float32)[..., 0].reshape(-1)
    edges = np.asarray(t.xb_bin_edges, np.float32)
    nb = len(edges) - 1
    b = np.clip(np.searchsorted(edges, xd, side="right") - 1, 0, nb - 1)
    tm = np.asarray(tn.tm_states).reshape(t.n_thin, t.n_sites)
    perm = (tm[t.xb_to_thin_id, np.clip(near, 0, t.n_sites - 1)] == 3) & np.asarray(t.xb_valid, bool)
    return dict(pos=np.concatenate([np.asarray(th.axial, np.float64).ravel(),
                                    np.asarray(tn.axial, np.float64).ravel()]),
                ls=float(ls), xb_bound=np.asarray(th.xb_bound_to, np.int32).ravel(),
                xb_key=(perm.astype(np.int64) * nb + b).astype(np.int32),
                xb_near=near.astype(np.int32), xb_xdist=xd,
                thin_bound=np.asarray(tn.bound_to, np.int32).ravel(),
                xb_state=np.asarray(th.xb_states, np.uint8).ravel(), tm=tm.astype(np.uint8).ravel())


def _from_record(rec: dict, topology, radial: Optional[float], prev: Optional[State] = None) -> State:
    """The engine's state record as a State.  ``radial`` is the spacing of the
    record's nearest-site update (the distances' radial column); None keeps
    ``prev``'s nearest-site fields (a solve does not touch them)."""
    t = topology
    shp = (t.n_thick, t.n_crowns, t.n_xb_per_crown)
    nk = t.n_thick * t.n_crowns
    if radial is None and prev is not None:
        near, dist = prev.thick.xb_nearest_bs, prev.thick.xb_distances
    else:
        near = rec["xb_near"].reshape(shp)
        dist = np.stack([rec["xb_xdist"], np.full(rec["xb_xdist"].shape, radial, np.float32)], -1).reshape(shp + (2,))
    return State(
        ThickState(rec["pos"][:nk].reshape(t.n_thick, t.n_crowns), rec["xb_state"].astype(np.int8).reshape(shp),
                   rec["xb_bound"].reshape(shp), near, dist),
        ThinState(rec["pos"][nk:].reshape(t.n_thin, t.n_sites), rec["tm"].astype(np.int8).reshape(t.n_thin, t.n_sites),
                  rec["thin_bound"].reshape(t.n_thin, t.n_sites)))


def _cfg(solver, dt, tolerance, n_newton_steps, n_cg_steps):
    from .engine import SolverConfig
    return SolverConfig(mode=solver, n_newton=int(n_newton_steps), n_cg=int(n_cg_steps),
                        tol=None if tolerance is None else float(tolerance), dt=float(dt))


def _uniforms(uniforms, topology):
    if uniforms is None:
        return None
    import torch
    u_tm, u_xb = uniforms
    S, X = topology.n_thin * topology.n_sites, topology.total_xbs
    u = np.concatenate([np.asarray(u_tm, np.float32).ravel(), np.asarray(u_xb, np.float32).ravel()])
    if u.size != S + X:
        raise ValueError(f"uniforms must hold {S} TM and {X} XB draws")
    return torch.from_numpy(u.reshape(1, 1, S + X)).cuda()


def _resolve(constants, drivers):
    row = _row(constants)
    c = {n: row[0, DYNAMIC_FIELDS.index(n)] for n in ("pCa", "z_line", "lattice_spacing")}
    if drivers is None:
        return row, c["pCa"], c["z_line"], c["lattice_spacing"]
    return (row, float(resolve_value(drivers.pCa, c["pCa"])), float(resolve_value(drivers.z_line, c["z_line"])),
            float(resolve_value(drivers.lattice_spacing, c["lattice_spacing"])))


def _step(state, constants, drivers, topology, rng_key, dt, K_lat, d_ref, tolerance, n_cg_steps,
          n_newton_steps, xb_subpop, tm_subpop, uniforms, solver, flags):
    from . import _lib
    row, pCa, z, ls = _resolve(constants, drivers)
    dynamic = K_lat is not None
    if dynamic and d_ref is None:
        raise ValueError("d_ref is required when K_lat is given")
    # the driver row: z_line, pCa and, with dynamic spacing, d_ref (the
    # kernel's Poisson reference is driver spacing x (z0 / z)^0 = d_ref);
    # the record's spacing is the kinetics spacing and the solve's first guess
    drv = np.array([[[z, pCa, float(d_ref) if dynamic else ls]]], np.float32)
    rrow = row.copy()
    for n, v in (("pCa", pCa), ("z_line", z), ("lattice_spacing", ls)):
        rrow[0, DYNAMIC_FIELDS.index(n)] = v          # resolved constants (with_drivers)
    pops = _populations(rrow, xb_subpop, tm_subpop, topology)
    bt = _batch(topology, rrow, drv, K_lat, pops)
    bt.set_state(0, _to_record(state, topology, ls))
    bt.seeds[0] = np.uint64(rng_key)
    bt._inst_cache.clear()
    bt.run(1, cfg=_cfg(solver, dt, tolerance, n_newton_steps, n_cg_steps),
           uniforms=_uniforms(uniforms, topology), launch_flags=flags)
    rec = bt.get_state(0)
    m = bt.metrics()[0, 0]
    kin_only = flags & _lib.MF_LAUNCH_KINETICS_ONLY
    new = _from_record(rec, topology, None if flags & _lib.MF_LAUNCH_SOLVE_ONLY else ls, prev=state)
    resolved = DynamicParams.from_row(rrow[0])
    if kin_only:
        return new, resolved, None, None, None
    return new, resolved, float(m[1]), float(rec["ls"]), int(m[51])


def kinetics_step(state: State, constants, drivers: Drivers, topology, rng_key, *, dt: float,
                  xb_subpop=None, tm_subpop=None, uniforms=None,
                  solver: str = "reference") -> Tuple[State, np.uint64, DynamicParams]:
    """The stochastic half of a step (timestep.py:65-129): drivers resolved,
    nearest sites updated, TM then XB transitions sampled; positions are
    left as they are.  Returns (state, new_rng_key, resolved_constants)."""
    from . import _lib
    new, resolved, *_ = _step(state, constants, drivers, topology, rng_key, dt, None, None, None, 6, 4,
                              xb_subpop, tm_subpop, uniforms, solver, _lib.MF_LAUNCH_KINETICS_ONLY)
    return new, split_key(rng_key), resolved


def solve_equilibrium(state: State, constants, topology, K_lat: float = None, d_ref: float = None,
                      tolerance: float = None, n_newton_steps: int = 16, n_cg_steps: int = 6,
                      precond_params=None, prefactored_precond=None, *,
                      solver: str = "reference") -> Tuple[State, float, float, int]:
    """Force balance of the given state (kernels/solver.py:663-771) at
    constants.z_line; fixed spacing constants.lattice_spacing, or with K_lat
    the spacing solved from that first guess against d_ref.  Returns
    (state, residual, new_lattice_spacing, n_iters)."""
    from . import _lib
    new, _, res, ls, it = _step(state, constants, None, topology, 0, 1.0, K_lat, d_ref, tolerance,
                                n_cg_steps, n_newton_steps, None, None, None, solver,
                                _lib.MF_LAUNCH_SOLVE_ONLY)
    return new, res, ls, it


def timestep(state: State, constants, drivers: Drivers, topology, rng_key, *, dt: float,
             K_lat=None, d_ref=None, solver_tol: Optional[float] = None, n_cg_steps: int = 6,
             n_newton_steps: int = 16, precond_params=None, prefactored_precond=None,
             xb_subpop=None, tm_subpop=None, uniforms=None,
             solver: str = "reference") -> Tuple[State, np.uint64, float, float, int]:
    """One step (timestep.py:136-199): kinetics_step then solve_equilibrium,
    in one launch.  Returns (state, new_rng_key, solver_residual, new_ls,
    n_iters)."""
    new, _, res, ls, it = _step(state, constants, drivers, topology, rng_key, dt, K_lat, d_ref, solver_tol,
                                n_cg_steps, n_newton_steps, xb_subpop, tm_subpop, uniforms, solver, 0)
    return new, split_key(rng_key), res, ls, it


def create_and_equilibrate(topology, constants, z_line: float, *, solver: str = "reference",
                           n_newton_steps: int = 4, n_cg_steps: int = 6) -> State:
    """realize_state at z_line, a nearest-site update, then the solve at the
    constants' own z_line and spacing (simulation.py:633-642; the reference's
    run() does not write its z_line into the constants)."""
    row = _row(constants)
    drv = np.array([[[z_line, row[0, DYNAMIC_FIELDS.index("pCa")],
                      row[0, DYNAMIC_FIELDS.index("lattice_spacing")]]]], np.float32)
    bt = _batch(topology, row, drv, None, None)
    bt.initialized[0] = False
    bt._inst_cache.clear()
    bt.run(0, cfg=_cfg(solver, 1.0, None, n_newton_steps, n_cg_steps))
    ls = float(row[0, DYNAMIC_FIELDS.index("lattice_spacing")])
    return _from_record(bt.get_state(0), topology, ls)
