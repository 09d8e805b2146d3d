"""Host mirror of the reference's ``run()`` / ``SimulationResult``
(multifil_jax/simulation.py:96-420, :754-1165) over the gfx950 engine.

``run()`` builds the same sweep grid as the reference:
* every sweepable input keeps a table of its variants; lists become
  Cartesian-product axes;
* ``replicates`` tiles the grid;
* the K_lat / nu and Poisson rules are applied (simulation.py:874-1022).

It then runs every instance in one persistent device launch per driver
segment and returns the metrics reshaped to ``grid + (replicates, time)``.

Differences from the reference, each forced by the engine:
* **Seeds.** Instance b draws Philox4x32-10 with seed
  ``(rng_seed << 32) | b`` instead of ``jax.random.split(PRNGKey(rng_seed))``.
* **No padding.** Nothing recompiles, so batches are not padded to power-of-two
  buckets, and ``unroll`` has no effect.
* **Solver mode.** ``solver`` selects the Newton mode:
  * "reference" (default): FP32, the reference's caps and tolerance floor;
  * "refined": FP64 residual to 1e-5 pN.
"""
from __future__ import annotations

import warnings
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from .params import DYNAMIC_FIELDS, DynamicParams, StaticParams

# The 52 metric keys of compute_all_metrics (metrics_fn.py:343-417), in the
# order the device writes them.
METRIC_NAMES: Tuple[str, ...] = (
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


class MetricsDict(dict):
    """dict with attribute access (state.py:126-136)."""

    def __getattr__(self, key):
        try:
            return self[key]
        except KeyError:
            raise AttributeError(key)


# ============================================================================
# SimulationResult (simulation.py:96-420)
# ============================================================================
class SimulationResult:
    """Result of run(): metric traces shaped (sweep_1, ..., replicates, time),
    the driver traces that produced them, and the grid's axis names and
    coordinates.  NumPy arrays throughout."""

    __slots__ = ('metrics', 'rng_key', 'z_line', 'pCa', 'lattice_spacing', 'final_state',
                 'metadata', 'dt', 'name', '_grid_shape', '_axis_names', 'coords',
                 'topology_config')

    def __init__(self, metrics, rng_key, z_line, pCa, lattice_spacing, final_state=None,
                 metadata=None, dt: float = 1.0, name: str = "", grid_shape=None,
                 axis_names=None, coords=None, topology_config=None):
        self.metrics = metrics if isinstance(metrics, MetricsDict) else MetricsDict(metrics)
        self.rng_key = rng_key
        self.z_line = z_line
        self.pCa = pCa
        self.lattice_spacing = lattice_spacing
        self.final_state = final_state
        self.metadata = metadata if metadata is not None else {}
        self.dt = float(dt)
        self.name = str(name)
        self._grid_shape = grid_shape
        self._axis_names = list(axis_names) if axis_names is not None else []
        self.coords = dict(coords) if coords is not None else {}
        self.topology_config = topology_config if topology_config is not None else {}

    @property
    def n_steps(self) -> int:
        return self.metrics['axial_force'].shape[-1]

    @property
    def replicate_axis(self) -> Optional[int]:
        return self._axis_names.index('replicates') if 'replicates' in self._axis_names else None

    @property
    def time(self) -> np.ndarray:
        return np.arange(self.n_steps) * self.dt

    @property
    def axial_force(self) -> np.ndarray:
        return self.metrics['axial_force']

    @property
    def mean_force(self) -> float:
        return float(np.mean(self.metrics['axial_force']))

    @property
    def steady_state_force(self) -> float:
        """Mean force over the last 20 % of the run."""
        n_avg = max(1, self.n_steps // 5)
        return float(np.mean(self.metrics['axial_force'][..., -n_avg:]))

    def _reduce_replicates(self, fn, suffix: str) -> 'SimulationResult':
        if self.replicate_axis is None:
            raise ValueError("No replicate axis to reduce.")
        new_metrics = MetricsDict({k: fn(v, axis=-2) for k, v in self.metrics.items()})
        force = new_metrics['axial_force']
        return SimulationResult(
            metrics=new_metrics, rng_key=self.rng_key,
            z_line=np.mean(self.z_line, axis=-2), pCa=np.mean(self.pCa, axis=-2),
            lattice_spacing=np.mean(self.lattice_spacing, axis=-2), final_state=None,
            metadata=self.metadata, dt=self.dt, name=self.name + suffix,
            grid_shape=force.shape[:-1] if force.ndim > 1 else None,
            axis_names=[n for n in self._axis_names if n != 'replicates'],
            coords={k: v for k, v in self.coords.items() if k != 'replicates'})

    def mean(self) -> 'SimulationResult':
        return self._reduce_replicates(np.mean, "_mean")

    def std(self) -> 'SimulationResult':
        return self._reduce_replicates(np.std, "_std")

    def __getitem__(self, key) -> 'SimulationResult':
        """Slice every array identically.  An integer drops its axis from the
        names and coords; a slice subsets the coords; anything else keeps the
        metadata as it was."""
        new_metrics = MetricsDict({k: v[key] for k, v in self.metrics.items()})
        force = new_metrics['axial_force']
        keys = key if isinstance(key, tuple) else (key,)
        names, coords = list(self._axis_names), dict(self.coords)
        opaque = False
        for pos, k in enumerate(keys):
            if pos >= len(self._axis_names):
                break
            name = self._axis_names[pos]
            if isinstance(k, (int, np.integer)):
                if name in names:
                    names.remove(name)
                coords.pop(name, None)
            elif isinstance(k, slice):
                if name in coords:
                    try:
                        coords[name] = list(coords[name])[k]
                    except Exception:
                        opaque = True
            else:
                opaque = True
        if opaque:
            names, coords = list(self._axis_names), dict(self.coords)
        return SimulationResult(
            metrics=new_metrics, rng_key=self.rng_key, z_line=self.z_line[key],
            pCa=self.pCa[key], lattice_spacing=self.lattice_spacing[key], final_state=None,
            metadata=self.metadata, dt=self.dt, name=self.name,
            grid_shape=force.shape[:-1] if force.ndim > 1 else None,
            axis_names=names, coords=coords)

    def sel(self, **kwargs) -> 'SimulationResult':
        """Coordinate-based slicing, e.g. ``result.sel(pCa=5.0)``."""
        idx = [slice(None)] * self.axial_force.ndim
        for axis_name, value in kwargs.items():
            if axis_name not in self.coords:
                raise ValueError(f"Unknown axis '{axis_name}'. Available: {list(self.coords.keys())}")
            coord_list = list(self.coords[axis_name])
            if value not in coord_list:
                raise ValueError(f"Value {value} not in {axis_name} coords: {coord_list}")
            idx[self._axis_names.index(axis_name)] = coord_list.index(value)
        out = self[tuple(idx)]
        out._axis_names = [n for n in self._axis_names if n not in kwargs]
        out.coords = {k: v for k, v in self.coords.items() if k not in kwargs}
        return out

    @classmethod
    def stack(cls, results: List['SimulationResult'], axis_name: str = 'structural',
              axis_values: Optional[List] = None) -> 'SimulationResult':
        """Stack results along a new leading named axis (structural sweeps)."""
        if not results:
            raise ValueError("Cannot stack empty list of results")
        keys = list(results[0].metrics.keys())
        coords = dict(results[0].coords)
        coords[axis_name] = axis_values if axis_values is not None else list(range(len(results)))
        metrics = MetricsDict({k: np.stack([r.metrics[k] for r in results]) for k in keys})
        return cls(metrics=metrics, rng_key=results[-1].rng_key,
                   z_line=np.stack([r.z_line for r in results]),
                   pCa=np.stack([r.pCa for r in results]),
                   lattice_spacing=np.stack([r.lattice_spacing for r in results]),
                   final_state=None, metadata=results[0].metadata, dt=results[0].dt,
                   name=results[0].name + "_stacked",
                   grid_shape=metrics['axial_force'].shape[:-1],
                   axis_names=[axis_name] + results[0]._axis_names, coords=coords,
                   topology_config=results[0].topology_config)

    def summary(self) -> str:
        lines = [
            f"SimulationResult: {self.name}",
            f"  Shape: {self.axial_force.shape}",
            f"  Duration: {self.n_steps * self.dt:.1f} ms ({self.n_steps} steps @ dt={self.dt}ms)",
            f"  Mean force: {self.mean_force:.2f} pN",
            f"  Steady-state force: {self.steady_state_force:.2f} pN",
            f"  Metrics: {list(self.metrics.keys()) if self.metrics else 'none'}",
            f"  Max solver residual: {float(np.max(self.metrics['solver_residual'])):.4f} pN",
        ]
        if self._axis_names:
            lines.append(f"  Grid axes: {self._axis_names}")
        if self.coords:
            lines.append(f"  Coords: {list(self.coords.keys())}")
        return '\n'.join(lines)

    def __repr__(self) -> str:
        return f"SimulationResult(name={self.name!r}, shape={self.axial_force.shape})"


# ============================================================================
# sweep grid (simulation.py:874-1022), engine-independent
# ============================================================================
class SweepPlan:
    """The batched inputs run() hands to the engine.

    Attributes: drivers (B, T, 3) float32 (z_line, pCa, lattice_spacing);
    params (B, 44) float32 rows; K_lat, nu (B,) float32 (K_lat already
    multiplied by n_thick); grid_shape, axis_names, coords; dynamic flag.
    """

    def __init__(self, n_steps, drivers, params, K_lat, nu, grid_shape, axis_names, coords,
                 dynamic, replicates):
        self.n_steps = n_steps
        self.drivers = drivers
        self.params = params
        self.K_lat = K_lat
        self.nu = nu
        self.grid_shape = grid_shape
        self.axis_names = axis_names
        self.coords = coords
        self.dynamic = dynamic
        self.replicates = replicates
        self.subpop = None       # subpopulation.ResolvedPopulations, when given

    @property
    def batch(self) -> int:
        return self.drivers.shape[0]


def plan_sweep(topology, duration_ms=1000.0, dt=1.0, pCa=4.5, z_line=900.0,
               lattice_spacing=14.0, K_lat=None, nu=0.0, dynamic_params=None,
               replicates=1, subpopulation=None) -> SweepPlan:
    """Build the Cartesian sweep exactly as run() does (simulation.py:874-1022)."""
    n_steps = int(duration_ms / dt)
    if n_steps < 1:
        raise ValueError(f"duration_ms / dt gives {n_steps} steps")
    if replicates < 1:
        raise ValueError("replicates must be >= 1")
    if dynamic_params is None or isinstance(dynamic_params, (dict, list)):
        base = DynamicParams()
    elif isinstance(dynamic_params, DynamicParams):
        base = dynamic_params
    else:
        raise ValueError(f"dynamic_params must be DynamicParams, dict, or list, "
                         f"got {type(dynamic_params)}")
    dynamic = K_lat is not None

    def as_trace(v):
        a = np.asarray(v, dtype=float)
        return np.full(n_steps, float(a)) if a.ndim == 0 else a

    sweep_axes = []
    wf_tables, wf_sweeps = {}, set()
    param_tables, param_axis = {}, {}
    for name, value in (('z_line', z_line), ('pCa', pCa), ('lattice_spacing', lattice_spacing)):
        if isinstance(value, list):
            if not value:
                raise ValueError(f"{name} sweep list is empty")
            wf_tables[name] = np.stack([as_trace(v) for v in value])
            if wf_tables[name].shape[1] != n_steps:
                raise ValueError(f"{name} trace length {wf_tables[name].shape[1]} != n_steps {n_steps}")
            is_trace = any(not isinstance(v, (int, float)) and np.asarray(v).ndim == 1
                           and len(np.asarray(v)) == n_steps for v in value)
            coords = [wf_tables[name][i] for i in range(len(value))] if is_trace else list(value)
            sweep_axes.append((name, coords))
            wf_sweeps.add(name)
        else:
            arr = np.asarray(value)
            if arr.ndim > 0 and arr.shape[0] != n_steps:
                raise ValueError(f"{name} array length {arr.shape[0]} != n_steps {n_steps}")
            wf_tables[name] = as_trace(value)[None, :]
    if isinstance(K_lat, list):
        sweep_axes.append(('K_lat', list(K_lat)))
    if isinstance(nu, list):
        sweep_axes.append(('nu', list(nu)))
    if isinstance(dynamic_params, list):
        if not dynamic_params:
            raise ValueError("dynamic_params candidate list is empty")
        sweep_axes.append(('candidates', list(range(len(dynamic_params)))))
        for name in DYNAMIC_FIELDS:
            param_tables[name] = np.array([np.float32(getattr(dp, name)) for dp in dynamic_params],
                                          np.float32)
            param_axis[name] = 'candidates'
    elif isinstance(dynamic_params, dict):
        for pname, pval in dynamic_params.items():
            if pname not in DYNAMIC_FIELDS:
                raise ValueError(f"Unknown dynamic_params field '{pname}'. "
                                 f"Valid: {list(DYNAMIC_FIELDS)}")
            if isinstance(pval, list):
                sweep_axes.append((pname, list(pval)))
                param_tables[pname] = np.asarray(pval, np.float32)
                param_axis[pname] = pname
    elif isinstance(dynamic_params, DynamicParams):
        for name in DYNAMIC_FIELDS:
            arr = np.asarray(getattr(dynamic_params, name))
            if arr.ndim > 0 and arr.shape[0] > 1:
                sweep_axes.append((name, arr.tolist()))
                param_tables[name] = arr.astype(np.float32)
                param_axis[name] = name
    # a subpopulation list is a candidate-style axis too (simulation.py:945-953)
    if isinstance(subpopulation, list):
        if isinstance(dynamic_params, list):
            raise ValueError("Cannot sweep a dynamic_params list and a subpopulation list "
                             "simultaneously (both are candidate-style axes).")
        if not subpopulation:
            raise ValueError("subpopulation list is empty")
        sweep_axes.append(('subpopulation', list(range(len(subpopulation)))))

    if sweep_axes:
        grids = np.meshgrid(*[np.arange(len(c)) for _, c in sweep_axes], indexing='ij')
        flat = {name: np.repeat(g.reshape(-1), replicates) for (name, _), g in zip(sweep_axes, grids)}
        batch = grids[0].size
    else:
        flat, batch = {}, 1
    total = batch * replicates
    grid_shape = tuple(len(c) for _, c in sweep_axes)
    axis_names = [n for n, _ in sweep_axes] + ['replicates', 'time']
    coords = {n: c for n, c in sweep_axes}
    coords['replicates'] = list(range(replicates))
    coords['time'] = (np.arange(n_steps) * dt).tolist()

    def waveform(name):
        t = wf_tables[name]
        return t[flat[name]] if name in wf_sweeps else np.broadcast_to(t[0], (total, n_steps))

    z_b, pCa_b, ls_b = waveform('z_line'), waveform('pCa'), waveform('lattice_spacing')
    if isinstance(K_lat, list):
        K_b = np.asarray(K_lat, np.float32)[flat['K_lat']]
    elif dynamic:
        K_b = np.full(total, float(K_lat), np.float32)
    else:
        K_b = np.zeros(total, np.float32)
    nu_b = (np.asarray(nu, np.float32)[flat['nu']] if isinstance(nu, list)
            else np.full(total, float(nu), np.float32))
    if dynamic:
        K_b = K_b * np.float32(topology.n_thick)
    ls_is_trace = not isinstance(lattice_spacing, list) and np.asarray(lattice_spacing).ndim > 0
    nonzero_nu = any(v != 0.0 for v in nu) if isinstance(nu, list) else float(nu) != 0.0
    if not dynamic and not ls_is_trace and nonzero_nu:
        # Poisson pre-computation: d = d0 (z0 / z)^nu (simulation.py:1005-1012)
        d0 = ls_b[:, 0:1]
        ls_b = d0 * (z_b[:, 0:1] / z_b) ** nu_b[:, None]

    def param(name):
        if name in param_axis:
            return param_tables[name][flat[param_axis[name]]]
        if isinstance(dynamic_params, dict) and name in dynamic_params:
            return np.full(total, float(dynamic_params[name]), np.float32)
        return np.full(total, float(getattr(base, name)), np.float32)

    params = np.stack([param(n) for n in DYNAMIC_FIELDS], axis=1).astype(np.float32)
    drivers = np.stack([z_b, pCa_b, ls_b], axis=-1).astype(np.float32)
    plan = SweepPlan(n_steps, np.ascontiguousarray(drivers), params, K_b, nu_b, grid_shape,
                     axis_names, coords, dynamic, replicates)
    if subpopulation is not None:
        from .subpopulation import resolve
        idx = flat['subpopulation'] if isinstance(subpopulation, list) else None
        plan.subpop = resolve(subpopulation, topology, total, idx)
    return plan


def _populations(plan: SweepPlan, Populations):
    """The engine's form of a resolved subpopulation: population rows,
    fractions and, for explicit modes, per-simulation labels."""
    from .subpopulation import population_params
    rp = plan.subpop
    prm = population_params(plan.params, rp)
    xl = tl = None
    if rp.mode == "explicit":
        xl = [m.astype(np.uint8) for m in rp.xb_mask]
        tl = [m.astype(np.uint8) for m in rp.tm_mask]
    return Populations(rp.mode, prm, rp.fractions, xl, tl, has_xb=rp.has_xb, has_tm=rp.has_tm)


# ============================================================================
# run() (simulation.py:754-1165)
# ============================================================================
def run(topology, duration_ms: float = 1000.0, dt: float = 1.0,
        pCa: Union[float, List[float], np.ndarray] = 4.5,
        z_line: Union[float, List[float], np.ndarray] = 900.0,
        lattice_spacing: Union[float, List[float], np.ndarray] = 14.0,
        K_lat: Union[float, List[float], None] = None,
        nu: Union[float, List[float]] = 0.0,
        dynamic_params: Union[DynamicParams, Dict, List, None] = None,
        static_params: Optional[StaticParams] = None,
        replicates: int = 1, rng_seed: int = 0, unroll: int = 1,
        minibatch_size: Optional[int] = "auto", verbose: bool = False,
        subpopulation=None, solver: str = "reference",
        block: Optional[int] = None) -> SimulationResult:
    """Run a batched half-sarcomere simulation on the GPU.

    Arguments and result shape follow the reference's run(); ``solver``
    ("reference" | "refined") and ``block`` (threads per workgroup) are
    engine options.  Instances whose sweep values share a lattice run in one
    launch; ``minibatch_size`` (an int) splits the batch into chunks to bound
    device memory, "auto" / None run everything at once.
    """
    from .engine import LatticePack, Batch, SolverConfig, Populations   # loads the device library
    static_params = static_params or StaticParams()
    plan = plan_sweep(topology, duration_ms, dt, pCa, z_line, lattice_spacing, K_lat, nu,
                      dynamic_params, replicates, subpopulation)
    pops = _populations(plan, Populations) if plan.subpop is not None else None
    B, T = plan.batch, plan.n_steps
    cfg = SolverConfig(mode=solver, n_newton=static_params.n_newton_steps,
                       n_cg=static_params.n_cg_steps, dt=dt)
    chunk = B if minibatch_size in (None, "auto") else max(1, int(minibatch_size))
    if verbose:
        print(f"Grid shape: {plan.grid_shape}, axes: {plan.axis_names}, total: {B}")
    pack = LatticePack([topology])
    seeds = (np.uint64(int(rng_seed) & 0xffffffff) << np.uint64(32)) | np.arange(B, dtype=np.uint64)
    out = np.empty((B, T, len(METRIC_NAMES)), np.float32)
    force = np.empty((B, T), np.float64)
    for s0 in range(0, B, chunk):
        s1 = min(B, s0 + chunk)
        bt = Batch(pack, np.zeros(s1 - s0, np.int32), plan.params[s0:s1], plan.drivers[s0:s1],
                   seeds=seeds[s0:s1],
                   dynamic=(plan.K_lat[s0:s1], plan.nu[s0:s1]) if plan.dynamic else None,
                   pops=pops.subset(np.arange(s0, s1)) if pops is not None else None)
        if block:
            bt.block = int(block)
        bt.run(T, cfg=cfg)
        out[s0:s1] = bt.metrics()
        force[s0:s1] = bt.force()
        del bt
    final_shape = plan.grid_shape + (replicates, T)
    metrics = MetricsDict({name: out[..., i].reshape(final_shape)
                           for i, name in enumerate(METRIC_NAMES)})
    metrics['axial_force'] = force.reshape(final_shape)
    max_res = float(np.max(metrics['solver_residual']))
    tol = static_params.solver_residual_tol
    if max_res > tol:
        warnings.warn(f"Solver max residual {max_res:.2f} pN exceeds threshold {tol} pN "
                      f"(set via StaticParams.solver_residual_tol)")
    if verbose:
        print(f"Result shape: {final_shape}; max solver residual: {max_res:.4f} pN")
    topo_cfg = {k: getattr(topology, k) for k in ('n_thick', 'n_crowns', 'n_thin', 'n_sites',
                                                  'n_titin', 'total_xbs', 'n_faces_per_thin')}
    if plan.dynamic:
        topo_cfg['K_lat'] = K_lat
        topo_cfg['K_lat_eff'] = (float(K_lat * topology.n_thick) if not isinstance(K_lat, list)
                                 else [float(k * topology.n_thick) for k in K_lat])
        topo_cfg['nu'] = nu
    return SimulationResult(
        metrics=metrics, rng_key=seeds[-1], z_line=plan.drivers[..., 0].reshape(final_shape),
        pCa=plan.drivers[..., 1].reshape(final_shape),
        lattice_spacing=plan.drivers[..., 2].reshape(final_shape), final_state=None,
        metadata={'master_seed': rng_seed, 'solver': solver}, dt=dt, name="run",
        grid_shape=plan.grid_shape, axis_names=plan.axis_names, coords=plan.coords,
        topology_config=topo_cfg)
