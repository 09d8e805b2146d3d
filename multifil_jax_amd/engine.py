"""Device-side batch engine: packs lattices and instances into flat HBM
buffers and drives the gfx950 kernels through the C ABI.

This replaces the reference's batch x time engine (multifil_jax/simulation.py:
584-747, ``_run_sim_kernel``): one persistent workgroup per instance loops
over timesteps on-chip instead of ``vmap(scan(timestep))``.

Layout in HBM (per launch):
  lattice blobs   float32 / int32 tables shared by every instance of a lattice
  params          (B, 44) float32, DynamicParams field order
  drivers         (B, T, 3) float32: z_line, pCa, lattice_spacing per step
  state           per-instance byte records (positions f64, bindings, states)
  tables          per-instance transition tables (cumulative P rows)
  metrics         (B, T, 52) float32, plus the FP64 M-line force (B, T)
PyTorch only provides device memory and the stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .params import N_PARAM_FIELDS, PARAM_INDEX

# driver columns and the parameter-row fields a NaN entry falls back to
DRIVER_FIELDS = ("z_line", "pCa", "lattice_spacing")

_STATE_FIELDS = ("pos", "ls", "xb_bound", "xb_key", "xb_near", "xb_xdist", "thin_bound", "xb_state", "tm")


def _al8(v: int) -> int:
    return (v + 7) & ~7


def state_layout(topo) -> dict:
    """Byte offsets of an instance state record (multifil_amd.h)."""
    N, X, S = topo.n_thick * topo.n_crowns + topo.n_thin * topo.n_sites, topo.total_xbs, topo.n_thin * topo.n_sites
    o = {"pos": 0}
    o["ls"] = 8 * N
    o["xb_bound"] = o["ls"] + 8
    o["xb_key"] = _al8(o["xb_bound"] + 4 * X)
    o["xb_near"] = _al8(o["xb_key"] + 4 * X)
    o["xb_xdist"] = _al8(o["xb_near"] + 4 * X)
    o["thin_bound"] = _al8(o["xb_xdist"] + 4 * X)
    o["xb_state"] = _al8(o["thin_bound"] + 4 * S)
    o["tm"] = o["xb_state"] + X
    o["bytes"] = _al8(o["tm"] + S)
    return o


@dataclass
class SolverConfig:
    """Solver settings of one launch.

    mode "reference": FP32 Newton-CG with the reference's caps (n_newton=4,
    n_cg=6; params.py:655-656) and tolerance floor (solver.py:714-718).
    mode "refined": the Newton residual in FP64 and the CG correction in FP32,
    iterated until max|F| <= refine_tol pN, or n_refine iterations (up to
    4 n_refine while each iteration still halves max|F|).
    """
    mode: str = "reference"
    n_newton: int = 4
    n_cg: int = 6
    n_refine: int = 8
    tol: Optional[float] = None
    refine_tol: float = 1e-5
    dt: float = 1.0
    cg_forcing: float = 0.0

    def ctype(self) -> _lib.MfSolverCfg:
        if self.mode not in ("reference", "refined"):
            raise ValueError(f"unknown solver mode {self.mode!r}")
        return _lib.MfSolverCfg(
            mode=_lib.MF_SOLVER_REFINED if self.mode == "refined" else _lib.MF_SOLVER_REFERENCE,
            n_newton=self.n_newton, n_cg=self.n_cg, n_refine=self.n_refine,
            tol=float("nan") if self.tol is None else float(self.tol),
            refine_tol=float(self.refine_tol), dt=float(self.dt), cg_forcing=float(self.cg_forcing))


class LatticePack:
    """Flat float32 / int32 blobs holding one or more lattices."""

    def __init__(self, topos: Sequence):
        f32: List[np.ndarray] = []
        i32: List[np.ndarray] = []
        self.topos = list(topos)
        descs = []
        nf = ni = 0

        def pf(a):
            nonlocal nf
            a = np.ascontiguousarray(a, dtype=np.float32).ravel()
            off = nf
            f32.append(a)
            nf += a.size
            return off

        def pi(a):
            nonlocal ni
            a = np.ascontiguousarray(a, dtype=np.int32).ravel()
            off = ni
            i32.append(a)
            ni += a.size
            return off

        for t in self.topos:
            if t.n_xb_bins > _lib.MF_MAX_BINS:
                raise ValueError(f"n_xb_bins {t.n_xb_bins} > {_lib.MF_MAX_BINS}")
            nfaces, width = t.n_faces_per_thin, t.max_sites_per_face
            # Candidate lists per (thin, face), padded exactly like the
            # reference's xb_to_site_indices (sarc_geometry.py:1475-1487).
            fs = np.zeros((t.n_thin, nfaces, width), np.int32)
            for i in range(t.n_thin):
                for f in range(nfaces):
                    n = int(t.n_sites_per_face[i, f])
                    first = int(t.face_to_sites[i, f, 0]) if n > 0 else 0
                    fs[i, f] = np.where(np.arange(width) < n, t.face_to_sites[i, f], first)
            slot = t.xb_to_thin_id.astype(np.int64) * nfaces + t.xb_to_thin_face
            # The face-table row of an XB must reproduce its own candidate row.
            if not np.array_equal(fs.reshape(-1, width)[slot], t.xb_to_site_indices):
                raise ValueError("xb_to_site_indices is not face_to_sites[thin, face]")
            # TM neighbours travel to the device as offsets along the filament
            k = np.arange(t.n_sites)[None, :]
            for nb_ in (t.tm_prev_neighbor, t.tm_next_neighbor):
                if np.any(np.abs(nb_ - k) > 127) or np.any(nb_ < 0) or np.any(nb_ >= t.n_sites):
                    raise ValueError("TM neighbours must lie within 127 sites on the same filament")
            co = t.crown_offsets.astype(np.float32)
            d = _lib.MfLattice(
                n_thick=t.n_thick, n_crowns=t.n_crowns, n_thin=t.n_thin, n_sites=t.n_sites,
                n_xb_per_crown=t.n_xb_per_crown, n_faces=nfaces, width=width, n_bins=t.n_xb_bins,
                crown_offsets=pf(co), crown_rests=pf(t.crown_rests),
                binding_offsets=pf(t.binding_offsets), binding_rests=pf(t.binding_rests),
                thin_rest_cum=pf(np.cumsum(t.binding_rests.astype(np.float32), axis=1, dtype=np.float32)),
                bin_edges=pf(t.xb_bin_edges), bin_centers=pf(t.xb_bin_centers),
                xb_thin=pi(t.xb_to_thin_id), xb_slot=pi(slot), xb_valid=pi(t.xb_valid.astype(np.int32)),
                face_sites=pi(fs), face_count=pi(t.n_sites_per_face),
                tm_prev=pi(t.tm_prev_neighbor), tm_next=pi(t.tm_next_neighbor),
                overlap_lo=float(np.float32(co.min()) - np.float32(13.0)),
                overlap_hi=float(np.float32(co.max()) + np.float32(13.0)))
            descs.append(d)
        self.f32 = np.concatenate(f32) if f32 else np.zeros(1, np.float32)
        self.i32 = np.concatenate(i32) if i32 else np.zeros(1, np.int32)
        self.desc_host = (_lib.MfLattice * len(descs))(*descs)
        self.desc_bytes = np.frombuffer(bytes(self.desc_host), dtype=np.uint8).copy()


def _torch():
    import torch
    return torch


def resolve_drivers(drivers: np.ndarray, params: np.ndarray) -> np.ndarray:
    """NaN driver entries take the instance's parameter-row value (the
    reference's sentinel: core/state.py:331-343 resolve_value, applied per
    step by timestep.py:97-102 and core/params.py:825-854 with_drivers).

    drivers: (B, T, 3) float32 (z_line, pCa, lattice_spacing); params:
    (B, 44) float32 rows.  Returns the array itself when it has no NaN, else
    a resolved copy.
    """
    nan = np.isnan(drivers)
    if not nan.any():
        return drivers
    const = params[:, [PARAM_INDEX[n] for n in DRIVER_FIELDS]].astype(np.float32)
    return np.where(nan, const[:, None, :], drivers).astype(np.float32)


def default_block(topos) -> int:
    """Workgroup size for a set of lattices: one wave per filament chain when
    the chains fit in 12 waves (768 threads, 168 VGPRs per lane), else 1024."""
    nch = max(t.n_thick + t.n_thin for t in topos)
    return 768 if nch <= 12 else 1024


class Populations:
    """Subpopulations of a batch in the engine's form (core/subpopulation.py;
    simulation.py:514-577, 645-677).

    mode: "mean_field" (one blended generator per kinetics, sum_k f_k Q_k)
        or "explicit" (every head / site carries a population label that
        selects its own transition rows).
    params: (B, K, 44) float32 population constants; row 0 of every
        instance is its wild-type row (the batch's params row), rows k > 0
        its scaled copies (only xb_* / tm_* fields may differ).
    fractions: (B, K) float32 mean-field weights.
    xb_labels: explicit: per instance a (X,) uint8 array of head labels;
    tm_labels: explicit: per instance a (S,) uint8 array of site labels.
    has_xb / has_tm: whether the populations change the XB / TM kinetics
        (any xb_* / tm_* field scaled); a kinetics they leave alone runs
        with the wild-type rows, as the reference passes subpop=None there.
    """

    def __init__(self, mode, params, fractions=None, xb_labels=None, tm_labels=None,
                 has_xb=True, has_tm=False):
        if mode not in ("mean_field", "explicit"):
            raise ValueError(f"unknown subpopulation mode {mode!r}")
        self.mode = mode
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        if self.params.ndim != 3 or self.params.shape[2] != N_PARAM_FIELDS:
            raise ValueError(f"population params must be (B, K, {N_PARAM_FIELDS})")
        self.B, self.K = self.params.shape[:2]
        if not 1 <= self.K <= _lib.MF_MAX_POPS:
            raise ValueError(f"1..{_lib.MF_MAX_POPS} populations supported, got {self.K}")
        self.fractions = (np.ascontiguousarray(fractions, dtype=np.float32) if fractions is not None
                          else np.full((self.B, self.K), 1.0 / self.K, np.float32))
        if self.fractions.shape != (self.B, self.K):
            raise ValueError(f"fractions must be ({self.B}, {self.K})")
        self.has_xb, self.has_tm = bool(has_xb), bool(has_tm)
        self.xb_labels = xb_labels
        self.tm_labels = tm_labels
        if mode == "explicit":
            for name, lab in (("xb_labels", xb_labels), ("tm_labels", tm_labels)):
                if lab is None or len(lab) != self.B:
                    raise ValueError(f"explicit subpopulations need {name} for every instance")
                for a in lab:
                    a = np.asarray(a)
                    if a.size and (a.min() < 0 or a.max() >= self.K):
                        raise ValueError(f"{name} must lie in [0, {self.K})")

    @property
    def flags(self) -> int:
        return (_lib.MF_POP_XB if self.has_xb else 0) | (_lib.MF_POP_TM if self.has_tm else 0)

    def table_sets(self):
        """(TM, XB) transition-table sets per instance."""
        ex = self.mode == "explicit"
        return (self.K if ex and self.has_tm else 1), (self.K if ex and self.has_xb else 1)

    def subset(self, idx) -> "Populations":
        idx = np.asarray(idx)
        pick = (lambda lab: [lab[i] for i in idx]) if self.mode == "explicit" else (lambda lab: None)
        return Populations(self.mode, self.params[idx], self.fractions[idx],
                           pick(self.xb_labels), pick(self.tm_labels), self.has_xb, self.has_tm)


class Batch:
    """B instances resident on one GPU, advanced by persistent launches.

    lattice_ids: (B,) index into ``pack.topos`` per instance.
    params: (B, 44) float32 parameter rows.
    drivers: (B, T, 3) float32 (z_line, pCa, lattice_spacing per step).
    seeds: (B,) uint64 Philox keys.
    dynamic: None, or (K_lat (B,), nu (B,)) for dynamic lattice spacing;
        K_lat is per instance and already multiplied by n_thick (as run()
        does, simulation.py:1002-1003); an instance with K_lat = NaN keeps a
        fixed spacing.
    lengths: optional (B,) steps per instance (<= T); an instance stops
        after its own length, its later metric rows stay zero.
    pops: optional Populations (subpopulations, one mode for the batch).
    """

    def __init__(self, pack: LatticePack, lattice_ids, params, drivers, seeds,
                 dynamic=None, device: str = "cuda", want_metrics: bool = True, lengths=None,
                 pops: Optional[Populations] = None):
        torch = _torch()
        lib = _lib.load()
        self.lib = lib
        self.pack = pack
        self.device = torch.device(device)
        self.lattice_ids = np.asarray(lattice_ids, dtype=np.int32)
        self.B = int(self.lattice_ids.size)
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.shape != (self.B, N_PARAM_FIELDS):
            raise ValueError(f"params must be ({self.B}, {N_PARAM_FIELDS}), got {params.shape}")
        drivers = np.ascontiguousarray(drivers, dtype=np.float32)
        if drivers.ndim != 3 or drivers.shape[0] != self.B or drivers.shape[2] != 3:
            raise ValueError(f"drivers must be ({self.B}, T, 3), got {drivers.shape}")
        self.T = drivers.shape[1]
        drivers = resolve_drivers(drivers, params)
        self.params_np = params
        self.drivers_np = drivers
        self.seeds = np.asarray(seeds, dtype=np.uint64)
        self.K_lat = np.full(self.B, np.nan, np.float32)
        self.nu = np.zeros(self.B, np.float32)
        if dynamic is not None:
            K, nu = dynamic
            self.K_lat[:] = np.broadcast_to(np.asarray(K, np.float32), (self.B,))
            self.nu[:] = np.broadcast_to(np.asarray(nu, np.float32), (self.B,))
        self.is_dynamic = np.isfinite(self.K_lat)
        self.lengths = (np.full(self.B, self.T, np.int64) if lengths is None
                        else np.asarray(lengths, np.int64))
        if self.lengths.shape != (self.B,) or np.any(self.lengths < 0) or np.any(self.lengths > self.T):
            raise ValueError(f"lengths must be ({self.B},) in [0, {self.T}]")

        topos = pack.topos
        self.layouts = [state_layout(topos[i]) for i in self.lattice_ids]
        sbytes = [lay["bytes"] for lay in self.layouts]
        for i, lay in zip(self.lattice_ids, self.layouts):
            t = topos[i]
            got = lib.mf_state_bytes(t.n_thick, t.n_crowns, t.n_thin, t.n_sites, t.n_xb_per_crown)
            if got != lay["bytes"]:
                raise RuntimeError("state layout mismatch between host and library")
        self.state_off = np.concatenate([[0], np.cumsum(sbytes)[:-1]]).astype(np.int64)
        self.pops = pops
        ktm, kxb = pops.table_sets() if pops is not None else (1, 1)
        if pops is not None:
            if pops.B != self.B:
                raise ValueError(f"populations cover {pops.B} instances, batch has {self.B}")
            if not np.array_equal(pops.params[:, 0], params):
                raise ValueError("population row 0 must equal the instance's parameter row")
        tfl = [int(lib.mf_table_floats_pop(topos[i].n_xb_bins, ktm, kxb)) for i in self.lattice_ids]
        self.table_off = np.concatenate([[0], np.cumsum(tfl)[:-1]]).astype(np.int64)
        self.label_off = np.zeros(self.B, np.int64)
        labels = None
        if pops is not None and pops.mode == "explicit":
            chunks = []
            for b, i in enumerate(self.lattice_ids):
                t = topos[i]
                X, S = t.total_xbs, t.n_thin * t.n_sites
                xl = np.asarray(pops.xb_labels[b], np.uint8).ravel()
                tl = np.asarray(pops.tm_labels[b], np.uint8).ravel()
                if xl.size != X or tl.size != S:
                    raise ValueError(f"instance {b}: labels must be ({X},) heads and ({S},) sites")
                chunks += [xl, tl]
            lens = np.array([c.size for c in chunks[0::2]]) + np.array([c.size for c in chunks[1::2]])
            self.label_off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
            labels = np.concatenate(chunks)

        dev = self.device
        self.d_lat = torch.from_numpy(pack.desc_bytes).to(dev)
        self.d_f32 = torch.from_numpy(pack.f32).to(dev)
        self.d_i32 = torch.from_numpy(pack.i32).to(dev)
        self.d_params = torch.from_numpy(params).to(dev)
        self.d_pop_params = self.d_pop_frac = self.d_labels = None
        if pops is not None:
            self.d_pop_params = torch.from_numpy(pops.params).to(dev)
            self.d_pop_frac = torch.from_numpy(pops.fractions).to(dev)
            if labels is not None:
                self.d_labels = torch.from_numpy(labels).to(dev)
        self.d_drivers = torch.from_numpy(drivers).to(dev)
        self.d_state = torch.zeros(int(sum(sbytes)), dtype=torch.uint8, device=dev)
        self.d_tables = torch.zeros(int(sum(tfl)), dtype=torch.float32, device=dev)
        self.d_metrics = (torch.zeros((self.B, self.T, _lib.MF_N_METRICS), dtype=torch.float32, device=dev)
                          if want_metrics else None)
        self.d_force = torch.zeros((self.B, self.T), dtype=torch.float64, device=dev)
        self._inst_cache = {}
        self.d_inst = None
        self.d_ws = None
        self._changes = None
        self.block = default_block(pack.topos)
        if self.is_dynamic.any() and self.block == 768:
            # the dynamic-spacing kernels need more registers: at 512 threads
            # (256 VGPRs per lane) a 2x2 dynamic step takes 155 us against
            # 178 us at 768, where a variant fits the lattices
            from .autotune import feasible_blocks
            if 512 in feasible_blocks(self, "refined"):
                self.block = 512
        self.initialized = np.zeros(self.B, dtype=bool)
        self._table_rows = None      # (B, 2) pCa / spacing the tables were built for
        self._table_dt = None

    # ------------------------------------------------------------------
    def _instances(self, n_steps: int, init_mask, s0: int = 0):
        """Device instance table for a launch over steps [s0, s0 + n_steps)
        (cached; the step offset itself travels in mf_launch.t_base)."""
        steps = np.clip(self.lengths - s0, 0, n_steps)
        key = (int(n_steps), init_mask.tobytes(), steps.tobytes())
        hit = self._inst_cache.get(key)
        if hit is None:
            arr = (_lib.MfInstance * self.B)()
            for b in range(self.B):
                arr[b].lattice = int(self.lattice_ids[b])
                arr[b].n_steps = int(steps[b])
                arr[b].t0 = 0
                arr[b].flags = (_lib.MF_INST_INIT if init_mask[b] else 0) | \
                    (_lib.MF_INST_DYNAMIC if self.is_dynamic[b] else 0)
                arr[b].K_lat = float(self.K_lat[b]) if self.is_dynamic[b] else 0.0
                arr[b].nu = float(self.nu[b])
                arr[b].seed = int(self.seeds[b])
                arr[b].state_off = int(self.state_off[b])
                arr[b].table_off = int(self.table_off[b])
                arr[b].label_off = int(self.label_off[b])
            buf = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
            hit = _torch().from_numpy(buf).to(self.device)
            self._inst_cache[key] = hit
        self.d_inst = hit

    def _launch(self, cfg: SolverConfig, uniforms=None, u_stride=0) -> _lib.MfLaunch:
        L = _lib.MfLaunch()
        L.lattices = self.d_lat.data_ptr()
        L.lat_f32 = self.d_f32.data_ptr()
        L.lat_i32 = self.d_i32.data_ptr()
        L.instances = self.d_inst.data_ptr()
        L.params = self.d_params.data_ptr()
        L.drivers = self.d_drivers.data_ptr()
        L.uniforms = uniforms.data_ptr() if uniforms is not None else None
        L.state = self.d_state.data_ptr()
        L.tables = self.d_tables.data_ptr()
        L.metrics = self.d_metrics.data_ptr() if self.d_metrics is not None else None
        L.force = self.d_force.data_ptr()
        L.n_lattices = len(self.pack.topos)
        L.B = self.B
        L.t_stride = self.T
        L.u_stride = int(u_stride)
        L.cfg = cfg.ctype()
        L.block = int(self.block)
        L.launch_flags = _lib.MF_LAUNCH_DYNAMIC if self.is_dynamic.any() else 0
        L.lattices_host = C.addressof(self.pack.desc_host)
        L.n_pops = 1
        if self.pops is not None:
            L.n_pops = self.pops.K
            L.pop_mode = _lib.MF_POP_MEAN_FIELD if self.pops.mode == "mean_field" else _lib.MF_POP_EXPLICIT
            L.pop_flags = self.pops.flags
            L.pop_params = self.d_pop_params.data_ptr()
            L.pop_fractions = self.d_pop_frac.data_ptr()
            L.pop_labels = self.d_labels.data_ptr() if self.d_labels is not None else None
        return L

    def _workspace(self, L):
        """Attach a device workspace when the chosen kernel keeps per-instance
        arrays in device memory (lattices too large for LDS)."""
        need = C.c_int64(0)
        _lib.check(self.lib.mf_workspace_bytes(C.byref(L), C.byref(need)), "mf_workspace_bytes")
        if need.value > 0:
            total = need.value * self.B
            if self.d_ws is None or self.d_ws.numel() < total:
                self.d_ws = _torch().empty(total, dtype=_torch().uint8, device=self.device)
            L.workspace = self.d_ws.data_ptr()
            L.workspace_stride = need.value

    def _stream(self):
        torch = _torch()
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def has_driver_changes(self) -> bool:
        """Whether any instance's pCa or spacing changes within the trace."""
        return len(self.segments(0, self.T)) > 1 if self.T > 0 else False

    def segments(self, t0: int, n: int) -> List[tuple]:
        """Split [t0, t0+n) where any instance's pCa or spacing changes: the
        transition tables are rebuilt between segments."""
        if self._changes is None:
            # steps k whose pCa / spacing differ from step k-1 (found once)
            pc = self.drivers_np[:, :, 1:3]
            self._changes = np.nonzero(np.any(np.any(pc[:, 1:] != pc[:, :-1], axis=2), axis=0))[0] + 1
        ch = self._changes[(self._changes > t0) & (self._changes < t0 + n)]
        cuts = [t0] + [int(c) for c in ch] + [t0 + n]
        return [(a, b - a) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]

    def _split(self, split: Optional[bool]) -> bool:
        # default: one launch per constant stretch for an all-fixed-spacing
        # batch (the tables kernel spreads each rebuild over the whole chip,
        # where an in-kernel rebuild runs on the instance's own workgroup),
        # in-kernel rebuilds when dynamic instances rebuild every step anyway
        return (not self.is_dynamic.any()) if split is None else bool(split)

    def prepare(self, n_steps: int, t0: int = 0, split: Optional[bool] = None):
        """Build and upload the instance tables that run(n_steps, t0) will
        use (and find the driver segments), so a timed run starts with every
        input already on the device."""
        init = ~self.initialized
        split = self._split(split)
        segs = self.segments(t0, n_steps) if (n_steps > 0 and split) else [(t0, n_steps)]
        for (s0, n) in segs:
            self._instances(n, init, s0)
            init = np.zeros_like(init)

    def run(self, n_steps: int, t0: int = 0, cfg: SolverConfig = None, uniforms=None,
            block: Optional[int] = None, sync: bool = True, kernel_events=None, noalloc: bool = False,
            launch_flags: int = 0, split: Optional[bool] = None):
        """Advance every instance from step t0 by n_steps (the first call
        equilibrates each instance before its first step, as the reference's
        create_and_equilibrate does).

        uniforms: optional device float32 tensor (B, T, S+X) of injected draws
        (TM sites first, then XBs) instead of the device Philox stream.
        kernel_events: optional (start, end) torch.cuda.Event pair recorded on
        the launch stream immediately around the step-kernel launches.
        noalloc: launch through mf_run_noalloc, whose large-lattice kernels
        use the library's static device scratch instead of a workspace this
        batch allocates.
        launch_flags: extra MF_LAUNCH_* bits (MF_LAUNCH_KINETICS_ONLY,
        MF_LAUNCH_SOLVE_ONLY: each step runs that half of the step alone).
        split: cut the run into one launch per stretch of constant pCa and
        spacing, with the tables rebuilt by mf_tables between launches
        (True), or run one launch in which each instance rebuilds its own
        tables when its drivers change (False, MF_LAUNCH_REBUILD).  Both give
        the same bits.  None: split unless the batch holds dynamic-spacing
        instances.
        """
        cfg = cfg or SolverConfig()
        split = self._split(split)
        if block is not None:
            self.block = int(block)
        if t0 < 0 or t0 + n_steps > self.T:
            raise ValueError(f"steps [{t0}, {t0 + n_steps}) outside the driver trace (T={self.T})")
        u_stride = 0
        if uniforms is not None:
            if tuple(uniforms.shape[:2]) != (self.B, self.T) or uniforms.dtype != _torch().float32:
                raise ValueError("uniforms must be float32 (B, T, S+X)")
            u_stride = uniforms.shape[2]
            need = max(t.n_thin * t.n_sites + t.total_xbs for t in self.pack.topos)
            if u_stride < need:
                raise ValueError(f"uniforms last dim {u_stride} < {need}")
            uniforms = uniforms.contiguous()
        stream = self._stream()
        changes = n_steps > 0 and len(self.segments(t0, n_steps)) > 1
        segs = self.segments(t0, n_steps) if (n_steps > 0 and split) else [(t0, n_steps)]
        for (s0, n) in segs:
            init = ~self.initialized
            self._instances(n, init, s0)
            L = self._launch(cfg, uniforms, u_stride)
            L.t_base = int(s0)
            L.launch_flags |= int(launch_flags)
            if changes and not split:
                # in-kernel table rebuilds run on the dynamic-spacing kernels
                # (fixed-spacing instances keep their fixed spacing there)
                L.launch_flags |= _lib.MF_LAUNCH_REBUILD | _lib.MF_LAUNCH_DYNAMIC
            rows = self.drivers_np[:, s0, 1:3]
            if self._table_rows is None or not np.array_equal(rows, self._table_rows) \
                    or cfg.dt != self._table_dt:
                _lib.check(self.lib.mf_tables(C.byref(L), stream), "mf_tables")
                self._table_rows = rows.copy()
                self._table_dt = cfg.dt
            if not noalloc:
                self._workspace(L)
            if kernel_events is not None and s0 == segs[0][0]:
                kernel_events[0].record(_torch().cuda.current_stream(self.device))
            if noalloc:
                _lib.check(self.lib.mf_run_noalloc(C.byref(L), stream), "mf_run_noalloc")
            else:
                _lib.check(self.lib.mf_run(C.byref(L), stream), "mf_run")
            if kernel_events is not None and s0 == segs[-1][0]:
                kernel_events[1].record(_torch().cuda.current_stream(self.device))
            self.initialized[:] = True
        if changes and not split:
            self._table_rows = None     # the kernel left the tables at other drivers
        if sync:
            _torch().cuda.synchronize(self.device)

    def update_inputs(self, params=None, drivers=None):
        """Replace the parameter rows and / or the driver trace in place
        (same shapes; NaN drivers resolved as at construction).  The next
        run rebuilds the transition tables."""
        torch = _torch()
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float32)
            if params.shape != self.params_np.shape:
                raise ValueError(f"params must be {self.params_np.shape}, got {params.shape}")
            self.params_np = params
            self.d_params.copy_(torch.from_numpy(params))
        if drivers is not None:
            drivers = np.ascontiguousarray(drivers, dtype=np.float32)
            if drivers.shape != self.drivers_np.shape:
                raise ValueError(f"drivers must be {self.drivers_np.shape}, got {drivers.shape}")
            self.drivers_np = resolve_drivers(drivers, self.params_np)
            self.d_drivers.copy_(torch.from_numpy(self.drivers_np))
            self._changes = None
        self._table_rows = None

    # ------------------------------------------------------------------
    def get_state(self, b: int) -> dict:
        """Copy one instance state to host numpy arrays."""
        t = self.pack.topos[self.lattice_ids[b]]
        lay = self.layouts[b]
        raw = self.d_state[self.state_off[b]:self.state_off[b] + lay["bytes"]].cpu().numpy()
        N, X, S = t.n_thick * t.n_crowns + t.n_thin * t.n_sites, t.total_xbs, t.n_thin * t.n_sites
        return {
            "pos": raw[lay["pos"]:lay["pos"] + 8 * N].view(np.float64).copy(),
            "ls": float(raw[lay["ls"]:lay["ls"] + 8].view(np.float64)[0]),
            "xb_bound": raw[lay["xb_bound"]:lay["xb_bound"] + 4 * X].view(np.int32).copy(),
            "xb_key": raw[lay["xb_key"]:lay["xb_key"] + 4 * X].view(np.int32).copy(),
            "xb_near": raw[lay["xb_near"]:lay["xb_near"] + 4 * X].view(np.int32).copy(),
            "xb_xdist": raw[lay["xb_xdist"]:lay["xb_xdist"] + 4 * X].view(np.float32).copy(),
            "thin_bound": raw[lay["thin_bound"]:lay["thin_bound"] + 4 * S].view(np.int32).copy(),
            "xb_state": raw[lay["xb_state"]:lay["xb_state"] + X].copy(),
            "tm": raw[lay["tm"]:lay["tm"] + S].copy(),
        }

    def set_state(self, b: int, st: dict):
        """Overwrite one instance state (marks it initialised)."""
        torch = _torch()
        t = self.pack.topos[self.lattice_ids[b]]
        lay = self.layouts[b]
        raw = np.zeros(lay["bytes"], np.uint8)
        N, X, S = t.n_thick * t.n_crowns + t.n_thin * t.n_sites, t.total_xbs, t.n_thin * t.n_sites

        def put(name, arr, dtype, n):
            a = np.ascontiguousarray(arr, dtype=dtype).ravel()
            if a.size != n:
                raise ValueError(f"{name}: expected {n} values, got {a.size}")
            raw[lay[name]:lay[name] + a.nbytes] = a.view(np.uint8)

        put("pos", st["pos"], np.float64, N)
        put("ls", [st["ls"]], np.float64, 1)
        put("xb_bound", st["xb_bound"], np.int32, X)
        put("xb_key", st["xb_key"], np.int32, X)
        # the last nearest-site update (the kernel writes these, never reads them)
        put("xb_near", st.get("xb_near", np.zeros(X, np.int32)), np.int32, X)
        put("xb_xdist", st.get("xb_xdist", np.zeros(X, np.float32)), np.float32, X)
        put("thin_bound", st["thin_bound"], np.int32, S)
        put("xb_state", st["xb_state"], np.uint8, X)
        put("tm", st["tm"], np.uint8, S)
        self.d_state[self.state_off[b]:self.state_off[b] + lay["bytes"]] = torch.from_numpy(raw).to(self.device)
        self.initialized[b] = True

    def metrics(self) -> np.ndarray:
        return self.d_metrics.cpu().numpy()

    def force(self) -> np.ndarray:
        return self.d_force.cpu().numpy()
