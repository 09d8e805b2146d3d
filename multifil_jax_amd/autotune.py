"""Autotuner for the step kernel's launch shape and CG cap.

Two knobs decide the speed of a batch:

* **Workgroup size.** It sets how many filament chains each wave owns, and
  the register budget per lane (768 threads: 168 VGPRs; 1024: 128).
* **CG iterations per Newton iteration** (refined mode only; reference mode
  keeps the reference's 6). Too few and Newton iterates more; too many and
  each iteration costs more.
* **CG forcing term** (refined mode only): 0 runs the capped CG count, a
  positive value stops CG once the correction is as accurate as the Newton
  step can use.

The tuner runs short trial launches of the same instances on a copy of the
batch. It tries every block size with a kernel variant for these lattices,
times each (block, n_cg) pair with HIP events, and keeps the fastest one
whose refined residual met its tolerance at every trial step. Choices are
cached for the process per (lattice shapes, the fixed solver settings,
dynamic spacing, subpopulation mode, batch size, a digest of the
conditions).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import hashlib
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .engine import Batch, SolverConfig, default_block

BLOCKS = (512, 768, 1024)
N_CG = (3, 4, 6, 8, 12)
CG_FORCING = (0.0, 0.03, 0.3)

_cache: Dict[tuple, Tuple[int, int, float]] = {}


def _key(batch: Batch, base: SolverConfig) -> tuple:
    """Cache key: lattice shapes, the solver settings the trials keep fixed,
    dynamic spacing, subpopulations, the batch size and a digest of the
    conditions (parameter rows, drivers, lengths, spacing stiffness)."""
    shapes = tuple(sorted({(t.n_thick, t.n_crowns, t.n_thin, t.n_sites) for t in batch.pack.topos}))
    pops = None if batch.pops is None else (batch.pops.mode, batch.pops.K, batch.pops.flags)
    h = hashlib.sha1()
    for a in (batch.lattice_ids, batch.params_np, batch.drivers_np, batch.lengths, batch.K_lat, batch.nu):
        h.update(np.ascontiguousarray(a).tobytes())
    fixed = (base.mode, base.n_newton, base.n_refine, base.tol, base.refine_tol, base.dt)
    return shapes, fixed, bool(batch.is_dynamic.any()), pops, batch.B, h.hexdigest()


def feasible_blocks(batch: Batch, mode: str) -> list:
    """Block sizes for which a kernel variant fits every lattice of the batch."""
    lib = _lib.load()
    out = []
    for blk in BLOCKS:
        L = _lib.MfLaunch()
        L.n_lattices = len(batch.pack.topos)
        L.B = batch.B
        L.block = blk
        L.cfg = SolverConfig(mode=mode).ctype()
        # driver traces run on the dynamic-spacing kernels (in-kernel rebuilds)
        L.launch_flags = _lib.MF_LAUNCH_DYNAMIC if (batch.is_dynamic.any() or batch.has_driver_changes()) else 0
        L.lattices_host = C.addressof(batch.pack.desc_host)
        L.n_pops = 1
        if batch.pops is not None:     # explicit populations widen the LDS TM rows
            L.n_pops = batch.pops.K
            L.pop_mode = _lib.MF_POP_MEAN_FIELD if batch.pops.mode == "mean_field" else _lib.MF_POP_EXPLICIT
            L.pop_flags = batch.pops.flags
            # the library checks that the population buffers a launch names exist
            L.pop_params = batch.d_pop_params.data_ptr()
            L.pop_fractions = batch.d_pop_frac.data_ptr()
            L.pop_labels = batch.d_labels.data_ptr() if batch.d_labels is not None else None
        lds = C.c_int64(0)
        if lib.mf_run_lds_bytes(C.byref(L), C.byref(lds)) == 0:
            out.append(blk)
    return out


def tune(batch: Batch, mode: str = "refined", trial_steps: int = 100, warmup_steps: int = 30,
         n_cg: Sequence[int] = N_CG, blocks: Optional[Sequence[int]] = None, max_instances: int = 256,
         verbose: bool = False, cg_forcing: Sequence[float] = CG_FORCING, repeats: int = 2,
         base: Optional[SolverConfig] = None):
    """Pick (block, n_cg, cg_forcing) for ``batch``; returns (block, SolverConfig).

    ``base`` carries the settings the trials keep (dt, Newton caps,
    tolerances); it defaults to ``SolverConfig(mode=mode)``.  The returned
    config is ``base`` with the tuned n_cg and cg_forcing.

    The trial copies up to ``max_instances`` instances (the longest-running
    ones when lengths differ), runs ``warmup_steps`` untimed steps past the
    activation transient, then times ``trial_steps`` steps; each candidate
    is timed ``repeats`` times (interleaved passes) and scored by its best.
    The batch itself is left untouched."""
    import torch
    base = SolverConfig(mode=mode) if base is None else dataclasses.replace(base, mode=mode)
    key = _key(batch, base)
    if key in _cache:
        blk, ncg, frc = _cache[key]
        return blk, dataclasses.replace(base, n_cg=ncg, cg_forcing=frc)
    if mode != "refined":
        # reference mode reproduces the reference's algorithm, whose CG count
        # (6, params.py:656) is part of it: only the launch shape is tuned
        n_cg = (6,)
        cg_forcing = (0.0,)
    W = min(warmup_steps, max(batch.T - 2, 1))
    T = min(W + trial_steps, batch.T)
    order = np.argsort(-batch.lengths, kind="stable")[:max_instances]
    K = batch.K_lat[order] if batch.is_dynamic.any() else None
    dyn = (K, batch.nu[order]) if K is not None else None
    blocks = list(blocks) if blocks is not None else feasible_blocks(batch, mode)
    best = None
    # every candidate is timed twice, in two interleaved passes, and scored
    # by its faster pass: candidates within a few percent of each other
    # (CG forcing on or off) are otherwise ranked by launch-to-launch noise
    times = {}
    cands = [(b_, n_, f_) for b_ in blocks for n_ in n_cg for f_ in cg_forcing]
    for blk, ncg, frc in cands * repeats:
        cfg = dataclasses.replace(base, n_cg=ncg, cg_forcing=frc)
        trial = Batch(batch.pack, batch.lattice_ids[order], batch.params_np[order],
                      batch.drivers_np[order, :T], seeds=batch.seeds[order], dynamic=dyn,
                      pops=batch.pops.subset(order) if batch.pops is not None else None)
        trial.block = blk
        trial.run(W, t0=0, cfg=cfg)                      # equilibrate + warm-up
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        trial.run(T - W, t0=W, cfg=cfg, sync=False)
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        resid = float(trial.metrics()[:, :T, 1].max())
        ok = mode != "refined" or resid <= cfg.refine_tol * (1 + 1e-6)
        if verbose:
            print(f"autotune block {blk} n_cg {ncg} forcing {frc}: {ms / (T - W) * 1e3:.1f} us/step, "
                  f"max residual {resid:.2e}{'' if ok else ' (rejected)'}")
        key_c = (blk, ncg, frc)
        if not ok:
            times[key_c] = None
        elif times.get(key_c, 0.0) is not None:
            times[key_c] = min(times.get(key_c, ms), ms)
        del trial
    for (blk, ncg, frc), ms in times.items():
        if ms is not None and (best is None or ms < best[0]):
            best = (ms, blk, ncg, frc)
    if best is None:
        blk, ncg, frc = default_block(batch.pack.topos), 6, 0.0
    else:
        _, blk, ncg, frc = best
    _cache[key] = (blk, ncg, frc)
    return blk, dataclasses.replace(base, n_cg=ncg, cg_forcing=frc)
