"""Benchmark: sarcomere-steps/s on the 256-condition 2x2 sweep.

Workload (batched sweep, the default): 256 half-sarcomere instances on a
2x2 lattice.
* pCa ~ U[4.0, 7.0] and z_line ~ U[850, 1150] nm (condition table drawn
  with numpy default_rng(20261203));
* dt = 1 ms, Philox seeds 0..255, fixed lattice spacing 14 nm;
* solver "refined": FP32 Newton-CG with an FP64 residual, to 1e-5 pN.

All inputs are resident in HBM before timing.  One bench "step" advances
every instance one timestep, so value = B * K / elapsed.

Also reported:
* the roofline record of the persistent step kernel (HBM), and its VALU
  issue rate against the VALU peak (the bound that applies);
* the plain per-condition GPU form (B = 1, one launch per timestep) on a
  bounded sample;
* the largest solver residual of the timed run;
* the CPU baseline: the C restatement of the step (oracle/cpu_port.c, FP32,
  the reference solver's caps) over the same sweep on up to 16 host
  threads.

The force-error bar of the metric (<= 1e-5 relative vs FP64) is checked
against the oracle by the GPU tests, not here.

Other workloads (``--workload``) print one line each:
* ``plain``: one condition (2x2, pCa 4.5, z 900 nm, 100 steps) persistent
  vs one launch per step;
* ``uneven``: 96 instances each of 2x2 / 3x3 / 4x4 / 6x6, durations
  U{200..1000}, pCa U[4.5, 6.5], dynamic lattice spacing (K_lat 5 pN/nm per
  thick filament, nu 0);
* ``sinusoid``: 4x4, 64 frequencies log-spaced 1-1000 Hz, amplitude 2.5 nm
  about z 1000 nm, pCa 4.5, 2000 steps at dt 0.1 ms, with the complex
  stiffness per frequency;
* ``lattices``: the step time of each lattice size (``--sizes``), 96
  instances of one lattice per launch, pCa U[4.5, 6.5], z 900 nm, fixed
  and / or dynamic spacing (``--spacing``), 30 timed steps after 10.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sarcomere-steps/s, 256-cond sweep: batched FP32 vs plain GPU, force err≤1e-5"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per sarcomere-step, 2x2 lattice, device RNG, 52 metrics:
# 2 * state (4N + X + 4X + S + 4S) + 12 (drivers) + 4*52 (metrics)
#   = 2 * 10432 + 12 + 208 = 21084 B.
BYTES_PER_SARC_STEP_2X2 = 21084
# Measured HBM traffic of the step kernel on this workload (rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE, scaled by a plain copy kernel), bytes per
# sarcomere-step.  The profile that holds it records its own command.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "run_kernel_2x2_sweep_traffic.md")
COPY_PROBE_MIB = 256
# Instruction mix of the same kernel (rocprofv3 SQ counters), and the VALU
# issue peak it is read against: one wave64 VALU instruction per CU per
# cycle (4 SIMD16 units x 4 cycles each), 256 CUs at 2.4 GHz.
SQ_PROFILE = os.path.join(ROOT, "profiles", "run_kernel_2x2_sweep_sq.md")
VALU_PEAK_GINST = 256 * 2.4


def measured_traffic_per_sarc_step():
    """Read 'traffic_bytes_per_sarcomere_step: <float>' from the committed
    traffic profile; None when it is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_sarcomere_step:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def valu_record(bt, K, ms_kernel):
    """The kernel against the compute side: VALU instructions issued per
    second (profiled count per wave-step x waves x steps / live kernel time)
    over the VALU issue peak.  The step is latency- and VALU-bound, not
    HBM-bound, so this is the fraction that says how close it runs."""
    per = measured_valu_per_wave_step()
    waves = bt.B * (bt.block // 64)
    if per is None or profiled_waves() != waves:
        return None             # the profile is of another launch shape
    achieved = per * waves * K / (ms_kernel / 1e3) / 1e9
    return {"insts_per_wave_step": per, "waves": waves, "achieved": achieved,
            "peak": VALU_PEAK_GINST, "unit": "G wave-instr/s", "frac": achieved / VALU_PEAK_GINST}


def profiled_waves():
    """Waves of the profiled dispatch ('Timed dispatch (K steps, W waves = ...')
    from the SQ profile."""
    import re
    try:
        with open(SQ_PROFILE) as f:
            m = re.search(r"Timed dispatch \(\d+ steps, (\d+) waves", f.read())
            return int(m.group(1)) if m else None
    except OSError:
        return None


def measured_valu_per_wave_step():
    """SQ_INSTS_VALU per wave per step from the committed SQ profile; None
    when it is absent."""
    try:
        with open(SQ_PROFILE) as f:
            for line in f:
                if line.startswith("| SQ_INSTS_VALU |"):
                    return float(line.strip().strip("|").split("|")[-1])
    except (OSError, ValueError):
        pass
    return None


def copy_probe(mib=COPY_PROBE_MIB, reps=3):
    """Plain device copy of a known size: the calibration dispatch for the
    FETCH_SIZE / WRITE_SIZE counters (reads and writes mib MiB each)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"bytes_read": mib << 20, "bytes_written": mib << 20, "reps": reps}


def conditions(B: int, T: int, seed: int = 20261203):
    rng = np.random.default_rng(seed)
    pCa = rng.uniform(4.0, 7.0, B).astype(np.float32)
    z = rng.uniform(850.0, 1150.0, B).astype(np.float32)
    drv = np.zeros((B, T, 3), np.float32)
    drv[:, :, 0] = z[:, None]
    drv[:, :, 1] = pCa[:, None]
    drv[:, :, 2] = 14.0
    return drv


def make_batch(B, T, lattice=2):
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch
    s, d = get_skeletal_params()
    topo = SarcTopology.create(lattice, lattice, s, d)
    drv = conditions(B, T)
    bt = Batch(LatticePack([topo]), np.zeros(B, np.int32), d.rows(B), drv,
               seeds=np.arange(B, dtype=np.uint64))
    return bt, topo, d, drv


def plain_gpu(cfg, B=256, n_steps=30, block=None):
    """The plain per-condition form: each of the sweep's B conditions alone
    (B = 1), one step-kernel launch per timestep, ``n_steps`` timed steps
    each after an untimed equilibration and first step."""
    import torch
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch
    s, d = get_skeletal_params()
    pack = LatticePack([SarcTopology.create(2, 2, s, d)])
    drv = conditions(B, n_steps + 1)
    t_all = 0.0
    for c in range(B):
        bt = Batch(pack, np.zeros(1, np.int32), d.rows(1), drv[c:c + 1], seeds=np.array([c], np.uint64))
        bt.run(1, t0=0, cfg=cfg, block=block)          # init + first step, untimed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(1, n_steps + 1):
            bt.run(1, t0=t, cfg=cfg, block=block, sync=False)
        torch.cuda.synchronize()
        t_all += time.perf_counter() - t0
    return B * n_steps / t_all, f"all {B} conditions x {n_steps} steps each, one launch per step"


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it: model name, vendor, family /
    model / stepping, the vector extensions the port can use, and the count
    of logical CPUs listed."""
    info, n = {}, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "processor":
                    n += 1
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    flags = set(info.get("flags", "").split())
    ext = "/".join(e for e in ("avx2", "fma", "avx512f", "avx512_bf16") if e in flags) or "none"
    return (f"{info.get('model name', 'unknown')} ({info.get('vendor_id', '?')} family {info.get('cpu family', '?')} "
            f"model {info.get('model', '?')} stepping {info.get('stepping', '?')}; {n} logical CPUs listed; "
            f"vector extensions {ext})")


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on
    (sched_getaffinity), capped by OMP_NUM_THREADS when the host sets it
    (the share of the machine this job is given)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    cap = int(env) if env and env.isdigit() and int(env) > 0 else aff
    return max(1, min(aff, cap)), aff


def cpu_baseline(B=256, T=1000, reps=3):
    """The C restatement of the reference algorithm (oracle/cpu_port.c: FP32,
    the reference solver's caps), OpenMP over conditions, on this host's
    cores: the sweep's own condition table (B conditions x T steps), one
    warm-up, best of ``reps``.  It is compiled here with -march=native
    (falling back to the portable in-tree build if that fails).  Transition
    tables are built before timing, as on the GPU.  The port computes the
    force, residual, Newton count and n_bound per step, not the other 48
    metrics: ``gpu_same_work`` times the GPU on that same work (reference
    solver, no metrics row)."""
    import tempfile
    from oracle import cpu_port
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    s, d = get_skeletal_params()
    topo = SarcTopology.create(2, 2, s, d)
    p = {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}
    threads, aff = cpu_threads()
    drv = conditions(B, T)
    seeds = np.arange(B, dtype=np.uint64)
    prep = cpu_port.prepare(topo, p, drv)
    try:
        lib_path, build = cpu_port.build_native(tempfile.mkdtemp(prefix="mfc_")), "gcc -O3 -march=native -fopenmp"
    except Exception as e:          # no compiler on the host: the portable build
        lib_path, build = None, f"in-tree x86-64-v2 build (native build failed: {type(e).__name__})"
    cpu_port.run(topo, p, drv[:threads, :20], seeds[:threads], threads=threads, lib_path=lib_path,
                 prepared={**prep, "tm": prep["tm"][:threads], "xb": prep["xb"][:threads]})
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        cpu_port.run(topo, p, drv, seeds, threads=threads, prepared=prep, lib_path=lib_path)
        el = time.perf_counter() - t0
        best = el if best is None else min(best, el)
    return {"value": B * T / best, "unit": "sarcomere-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "affinity_cpus": aff, "cpu": cpu_model(), "build": build,
            "sample": f"the full sweep ({B} conditions x {T} steps), C restatement (oracle/cpu_port.c, FP32, "
                      f"reference solver caps 4 x 6, force / residual / iterations / n_bound per step), "
                      f"OpenMP {threads} threads ({aff} CPUs in the affinity mask, OMP_NUM_THREADS "
                      f"{os.environ.get('OMP_NUM_THREADS', 'unset')}), best of {reps}"}


def gpu_same_work(B=256, T=1000, W=20):
    """The GPU on the CPU port's work: the same sweep, reference solver,
    force only (no 52-metric row); sarcomere-steps/s over T steps."""
    from multifil_jax_amd.engine import SolverConfig
    import torch
    bt, *_ = make_batch(B, W + T)
    bt.d_metrics = None
    cfg = SolverConfig(mode="reference")
    bt.run(W, t0=0, cfg=cfg)
    bt.prepare(T, t0=W)
    torch.cuda.synchronize()
    ms = _timed(bt, T, cfg, t0=W)
    return B * T / (ms / 1e3)


def _timed(bt, n_steps, cfg, t0=0):
    """Run n_steps from t0 between two events on the launch stream; ms."""
    import torch
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    bt.run(n_steps, t0=t0, cfg=cfg, sync=False)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def workload_plain(cfg):
    """One condition (2x2, pCa 4.5, z 900 nm, 14 nm, seed 0), 100 steps:
    one persistent launch vs one launch per step."""
    import torch
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch
    s, d = get_skeletal_params()
    topo = SarcTopology.create(2, 2, s, d)
    T = 100
    drv = np.zeros((1, T + 1, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = 900.0, 4.5, 14.0
    res = {}
    for form in ("persistent", "launch_per_step"):
        bt = Batch(LatticePack([topo]), np.zeros(1, np.int32), d.rows(1), drv, seeds=np.zeros(1, np.uint64))
        bt.run(1, t0=0, cfg=cfg)                       # equilibrate + step 0, untimed
        if form == "persistent":
            ms = _timed(bt, T, cfg, t0=1)
        else:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for t in range(1, T + 1):
                bt.run(1, t0=t, cfg=cfg, sync=False)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
        res[form] = T / (ms / 1e3)
    return {"metric": "sarcomere-steps/s, single condition 2x2 (plain method)",
            "value": res["persistent"], "unit": "sarcomere-steps/s", "n_gpus": 1, "steps": T,
            "higher_is_better": True, "dtype": "f32+f64refine" if cfg.mode == "refined" else "f32",
            "data": "synthetic",
            "config": {"workload": "plain: 2x2, pCa 4.5, z_line 900 nm, spacing 14 nm, dt 1 ms, 100 steps, seed 0",
                       "solver": cfg.mode},
            "launch_per_step": res["launch_per_step"]}


def uneven_conditions(seed=20261203, per_lattice=96, lattices=(2, 3, 4, 6)):
    rng = np.random.default_rng(seed)
    n = per_lattice * len(lattices)
    lat = np.repeat(np.arange(len(lattices)), per_lattice)
    lengths = rng.integers(200, 1001, n)
    pCa = rng.uniform(4.5, 6.5, n).astype(np.float32)
    return lat, lengths, pCa


def uneven_batches(cfg, block=None, autotune=True):
    """The uneven workload's two launches: 96 instances each of 2x2 / 3x3 /
    4x4 / 6x6, lengths U{200..1000}, dynamic spacing.  The LDS-resident
    lattices share one batch; 6x6 (device workspace) gets its own, run first
    on its own stream.  Returns (batches, T, total sarcomere-steps); each
    batch carries ``tuned_cfg``."""
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch
    s, d = get_skeletal_params()
    sizes = (2, 3, 4, 6)
    topos = [SarcTopology.create(n, n, s, d) for n in sizes]
    lat, lengths, pCa = uneven_conditions()
    T = int(lengths.max())
    # the 6x6 launch goes first: its instances are the longest jobs, and the
    # other launch's 288 workgroups would otherwise take every CU before it
    groups = [np.nonzero(lat == 3)[0], np.nonzero(lat < 3)[0]]
    batches = []
    for gi, idx in enumerate(groups):
        # longest jobs first: the dispatcher hands out workgroups in order
        cost = lengths[idx] * np.array([t.n_thick * t.n_crowns + t.n_thin * t.n_sites for t in topos])[lat[idx]]
        idx = idx[np.argsort(-cost, kind="stable")]
        gtopos = [topos[i] for i in sorted(set(lat[idx]))]
        remap = {i: k for k, i in enumerate(sorted(set(lat[idx])))}
        ids = np.array([remap[i] for i in lat[idx]], np.int32)
        B = len(idx)
        drv = np.zeros((B, T, 3), np.float32)
        drv[..., 0], drv[..., 2] = 900.0, 14.0
        drv[..., 1] = pCa[idx][:, None]
        K = np.array([5.0 * topos[i].n_thick for i in lat[idx]], np.float32)
        bt = Batch(LatticePack(gtopos), ids, d.rows(B), drv, seeds=idx.astype(np.uint64),
                   dynamic=(K, np.zeros(B, np.float32)), lengths=lengths[idx])
        bt.tuned_cfg = cfg
        if block:
            bt.block = block
        elif autotune:
            from multifil_jax_amd.autotune import tune
            bt.block, bt.tuned_cfg = tune(bt, mode=cfg.mode, base=cfg)
        batches.append(bt)
    return batches, T, int(lengths.sum())


def run_uneven(batches, T):
    """Both launches, concurrently on two streams; wall seconds."""
    import torch
    streams = [torch.cuda.Stream() for _ in batches]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for bt, st in zip(batches, streams):
        with torch.cuda.stream(st):
            bt.run(T, cfg=bt.tuned_cfg, sync=False)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def workload_uneven(cfg, block=None, autotune=True):
    """96 instances each of 2x2 / 3x3 / 4x4 / 6x6, lengths U{200..1000},
    dynamic spacing, in two concurrent launches."""
    batches, T, total = uneven_batches(cfg, block, autotune)
    el = run_uneven(batches, T)
    worst = max(float(bt.metrics()[:, :, 1].max()) for bt in batches)
    return {"metric": "sarcomere-steps/s, uneven batch 384 x {2x2,3x3,4x4,6x6}, dynamic spacing",
            "value": total / el, "unit": "sarcomere-steps/s", "n_gpus": 1, "steps": T,
            "higher_is_better": True, "dtype": "f32+f64refine" if cfg.mode == "refined" else "f32",
            "data": "synthetic", "sarcomere_steps": total, "seconds": el,
            "max_solver_residual_pN": worst,
            "config": {"workload": "uneven: 96 each of 2x2/3x3/4x4/6x6, lengths U{200..1000}, pCa U[4.5,6.5], "
                                   "z_line 900 nm, dynamic spacing K_lat 5 pN/nm per thick, nu 0, d0 14 nm, "
                                   "dt 1 ms; default_rng(20261203)",
                       "solver": cfg.mode, "launches": 2,
                       "groups": [{"block": bt.block, "n_cg": bt.tuned_cfg.n_cg,
                                   "cg_forcing": bt.tuned_cfg.cg_forcing,
                                   "instances": bt.B} for bt in batches]}}


def workload_sinusoid(cfg):
    """4x4, 64 frequencies logspace(0, 3) Hz, 2.5 nm about 1000 nm, pCa 4.5,
    2000 steps at dt 0.1 ms; complex stiffness per frequency."""
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    from multifil_jax_amd.analysis import sinusoid_traces, complex_stiffness
    s, d = get_skeletal_params()
    topo = SarcTopology.create(4, 4, s, d)
    freqs = np.logspace(0, 3, 64)
    T, dt = 2000, 0.1
    drv = np.zeros((64, T, 3), np.float32)
    drv[..., 0] = sinusoid_traces(freqs, 1000.0, 2.5, dt, T)
    drv[..., 1], drv[..., 2] = 4.5, 14.0
    c = SolverConfig(mode=cfg.mode, n_cg=cfg.n_cg, dt=dt)
    bt = Batch(LatticePack([topo]), np.zeros(64, np.int32), d.rows(64), drv, seeds=np.arange(64, dtype=np.uint64))
    ms = _timed(bt, T, c)
    st = complex_stiffness(bt.force()[:, None, :], freqs, dt, 1000.0, 2.5, settling_ms=20.0)
    ok = ~st["too_short"]
    return {"metric": "sarcomere-steps/s, sinusoidal 4x4 x 64 frequencies",
            "value": 64 * T / (ms / 1e3), "unit": "sarcomere-steps/s", "n_gpus": 1, "steps": T,
            "higher_is_better": True, "dtype": "f32+f64refine" if cfg.mode == "refined" else "f32",
            "data": "synthetic", "max_solver_residual_pN": float(bt.metrics()[:, :, 1].max()),
            "config": {"workload": "sinusoid: 4x4, 64 freqs logspace(0,3) Hz, amplitude 2.5 nm about z 1000 nm, "
                                   "pCa 4.5, spacing 14 nm, dt 0.1 ms, 2000 steps",
                       "solver": cfg.mode, "settling_ms": 20.0},
            "stiffness": {"freq_hz": [round(float(f), 4) for f in freqs[ok]],
                          "elastic_pN_per_nm": [round(float(v), 5) for v in st["elastic"][ok]],
                          "viscous_pN_per_nm": [round(float(v), 5) for v in st["viscous"][ok]],
                          "too_short_hz": [round(float(f), 4) for f in freqs[~ok]]}}


def workload_lattices(sizes, spacing):
    """Per-lattice step time: 96 instances of one lattice per launch, pCa
    U[4.5, 6.5] (default_rng(5)), z 900 nm, spacing 14 nm, refined solver
    (CG cap 6, forcing 0.003), 10 untimed steps then 30 timed; dynamic
    spacing with K_lat 5 pN/nm per thick filament, nu 0."""
    import torch
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    s, d = get_skeletal_params()
    B, W, K = 96, 10, 30
    c = SolverConfig(mode="refined", n_cg=6, cg_forcing=0.003)
    pca = np.random.default_rng(5).uniform(4.5, 6.5, B).astype(np.float32)
    modes = {"fixed": (False,), "dynamic": (True,), "both": (False, True)}[spacing]
    rows = []
    for n in sizes:
        topo = SarcTopology.create(n, n, s, d)
        for dyn in modes:
            drv = np.zeros((B, W + K, 3), np.float32)
            drv[..., 0], drv[..., 2] = 900.0, 14.0
            drv[..., 1] = pca[:, None]
            kw = {"dynamic": (np.full(B, 5.0 * topo.n_thick, np.float32), np.zeros(B, np.float32))} if dyn else {}
            bt = Batch(LatticePack([topo]), np.zeros(B, np.int32), d.rows(B), drv,
                       seeds=np.arange(B, dtype=np.uint64), **kw)
            try:
                bt.run(W, cfg=c)
            except RuntimeError as e:   # a lattice / spacing the engine refuses (before any launch)
                rows.append({"lattice": f"{n}x{n}", "spacing": "dynamic" if dyn else "fixed", "refused": str(e)})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
                continue
            bt.prepare(K, t0=W)
            torch.cuda.synchronize()
            ms = _timed(bt, K, c, t0=W)
            m = bt.metrics()
            rows.append({"lattice": f"{n}x{n}", "spacing": "dynamic" if dyn else "fixed", "block": bt.block,
                         "us_per_step": round(1e3 * ms / K, 2),
                         "newton_mean": round(float(m[:, W:, 51].mean()), 3),
                         "max_solver_residual_pN": float(m[:, W:, 1].max())})
            print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    return {"metric": "us per step of 96 instances, per lattice", "unit": "us", "higher_is_better": False,
            "n_gpus": 1, "steps": K, "data": "synthetic",
            "config": {"workload": "lattices: 96 instances of one lattice per launch, pCa U[4.5,6.5], z 900 nm, "
                                   "spacing 14 nm (dynamic: K_lat 5 pN/nm per thick, nu 0), dt 1 ms, "
                                   "10 + 30 steps", "solver": "refined", "n_cg": 6, "cg_forcing": 0.003},
            "results": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--mode", default="refined", choices=["refined", "reference"])
    ap.add_argument("--block", type=int, default=None,
                    help="threads per workgroup (default: the engine's choice for the lattice)")
    ap.add_argument("--n-cg", type=int, default=6, help="CG iterations per Newton step")
    ap.add_argument("--no-extras", action="store_true", help="skip plain/oracle/cpu legs")
    ap.add_argument("--workload", default="sweep", choices=["sweep", "plain", "uneven", "sinusoid", "lattices"])
    ap.add_argument("--sizes", default="2,3,4,6", help="lattices workload: n of the n x n lattices")
    ap.add_argument("--spacing", default="both", choices=["fixed", "dynamic", "both"],
                    help="lattices workload: spacing modes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the block-size / CG-cap autotuner (sweep and uneven workloads)")
    ap.add_argument("--copy-probe", action="store_true",
                    help=f"after timing, copy {COPY_PROBE_MIB} MiB on the device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is "
                                   "designed (per-instance partition) but not built"}))
        sys.exit(2)
    import torch
    from multifil_jax_amd.engine import SolverConfig
    cfg = SolverConfig(mode=args.mode, n_cg=args.n_cg)
    if args.workload != "sweep":
        fn = {"plain": workload_plain,
              "uneven": lambda c: workload_uneven(c, args.block, not args.no_autotune),
              "sinusoid": workload_sinusoid,
              "lattices": lambda c: workload_lattices([int(x) for x in args.sizes.split(",")], args.spacing)}[
                  args.workload]
        print(json.dumps(fn(cfg)))
        return
    B, K, W = args.batch, args.steps, args.warmup
    bt, topo, d, drv = make_batch(B, W + K)
    tuned = None
    if args.block:
        bt.block = args.block
    elif not args.no_autotune:
        from multifil_jax_amd.autotune import tune
        bt.block, cfg = tune(bt, mode=args.mode, base=cfg)
        tuned = {"block": bt.block, "n_cg": cfg.n_cg}
    # warmup: equilibration + W steps (also builds the transition tables)
    bt.run(W, t0=0, cfg=cfg)
    torch.cuda.synchronize()
    bt.prepare(K, t0=W)                  # instance tables on the device before timing
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    k0 = torch.cuda.Event(enable_timing=True)
    k1 = torch.cuda.Event(enable_timing=True)
    t_host0 = time.perf_counter()
    e0.record(stream)
    # one persistent step-kernel launch covers all K steps of all B instances
    bt.run(K, t0=W, cfg=cfg, sync=False, kernel_events=(k0, k1))
    e1.record(stream)
    torch.cuda.synchronize()
    t_host = time.perf_counter() - t_host0
    ms = e0.elapsed_time(e1)
    ms_kernel = k0.elapsed_time(k1)
    value = B * K / (ms / 1e3)
    bytes_launch = BYTES_PER_SARC_STEP_2X2 * B * K
    achieved = bytes_launch / (ms_kernel / 1e3) / 1e9
    tps = measured_traffic_per_sarc_step()
    traffic = None if tps is None else tps * B * K
    out = {
        "metric": METRIC, "value": value, "unit": "sarcomere-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": ms / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64refine" if args.mode == "refined" else "f32",
        "data": "synthetic",
        "config": {"workload": "batched sweep: 256 conditions on a 2x2 lattice, pCa U[4,7], "
                               "z_line U[850,1150] nm, dt 1 ms, seeds 0..255, fixed spacing 14 nm",
                   "lattice": "2x2", "batch": B, "solver": args.mode, "block": bt.block,
                   "n_cg": cfg.n_cg, "cg_forcing": cfg.cg_forcing, "autotuned": tuned is not None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "run_kernel", "kernel_ms": ms_kernel, "launches": 1},
        "host_s": t_host,
        "max_solver_residual_pN": float(bt.metrics()[:, W:W + K, 1].max()),
        "valu": valu_record(bt, K, ms_kernel),
    }
    if args.copy_probe:
        out["copy_probe"] = copy_probe()
    if not args.no_extras:
        pv, psample = plain_gpu(cfg)
        out["plain_gpu"] = {"value": pv, "unit": "sarcomere-steps/s", "sample": psample}
        out["batched_over_plain"] = value / pv
        out["cpu_baseline"] = cpu_baseline()
        g = gpu_same_work()
        out["cpu_baseline"]["gpu_same_work"] = {
            "value": g, "unit": "sarcomere-steps/s",
            "what": "the GPU on the port's work: same sweep, reference solver, force only",
            "ratio": g / out["cpu_baseline"]["value"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
