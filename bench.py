"""Benchmark: sarcomere-steps/s on the 256-condition 2x2 sweep.

Workload (batched sweep): 256 half-sarcomere instances on a 2x2 lattice,
pCa ~ U[4.0, 7.0], z_line ~ U[850, 1150] nm (condition table drawn with
numpy default_rng(20261203)), dt = 1 ms, Philox seeds 0..255, fixed lattice
spacing 14 nm, solver "refined" (FP32 Newton-CG with an FP64 residual).
All inputs are resident in HBM before timing.  One bench "step" advances
every instance one timestep, so value = B * K / elapsed.

Also reported: the roofline record of the persistent step kernel, the plain
per-condition GPU form (B = 1, one launch per timestep, no graph) on a
bounded sample, the FP64-oracle force error on a sample, and the CPU
baseline (the NumPy FP64 oracle on a bounded sample).
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


def plain_gpu(cfg, n_cond=8, n_steps=100, block=None):
    """B = 1 per condition, one persistent-kernel launch per timestep."""
    import torch
    total = 0
    t_all = 0.0
    for c in range(n_cond):
        bt, *_ = make_batch(1, n_steps + 1)
        bt.run(1, t0=0, cfg=cfg, block=block)          # init + first step, untimed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(1, n_steps + 1):
            bt.run(1, t0=t, cfg=cfg, block=block, sync=False)
        torch.cuda.synchronize()
        t_all += time.perf_counter() - t0
        total += n_steps
    return total / t_all, f"{n_cond} conditions x {n_steps} steps, one launch per step"


def force_error_sample(cfg, n_inst=2, n_steps=20):
    """Max relative M-line force error of the device vs the FP64 oracle, same draws."""
    from oracle import mf_oracle as mo
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    bt, topo, d, drv = make_batch(n_inst, n_steps)
    bt.run(n_steps, cfg=cfg)
    fg = bt.force()
    p = {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}
    worst = 0.0
    for b in range(n_inst):
        _, out = mo.run(topo, p, drv[b, :, 0], drv[b, :, 1], drv[b, :, 2], 1.0, b,
                        mo.SolverConfig(mode="exact" if cfg.mode == "refined" else "reference"))
        worst = max(worst, float(np.max(np.abs(fg[b] - out[:, 0]) / np.abs(out[:, 0]))))
    return worst


def cpu_baseline(seconds=10.0):
    """FP64 NumPy oracle, single core, bounded sample of the same workload."""
    from oracle import mf_oracle as mo
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    from multifil_jax_amd.topology import SarcTopology
    from multifil_jax_amd.params import get_skeletal_params
    s, d = get_skeletal_params()
    topo = SarcTopology.create(2, 2, s, d)
    p = {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}
    drv = conditions(256, 1)
    steps = 0
    t0 = time.perf_counter()
    c = 0
    while time.perf_counter() - t0 < seconds:
        n = 50
        mo.run(topo, p, np.full(n, drv[c, 0, 0]), np.full(n, drv[c, 0, 1]), np.full(n, 14.0),
               1.0, c, mo.SolverConfig(mode="exact"))
        steps += n
        c += 1
    el = time.perf_counter() - t0
    return {"value": steps / el, "unit": "sarcomere-steps/s", "cores": 1, "kind": "port",
            "sample": f"{c} conditions x 50 steps (incl. equilibration), FP64 NumPy oracle, exact solve"}


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
    B, K, W = args.batch, args.steps, args.warmup
    bt, topo, d, drv = make_batch(B, W + K)
    if args.block:
        bt.block = args.block
    # warmup: equilibration + W steps (also builds the transition tables)
    bt.run(W, t0=0, cfg=cfg)
    torch.cuda.synchronize()
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
                   "n_cg": args.n_cg},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "run_kernel", "kernel_ms": ms_kernel, "launches": 1},
        "host_s": t_host,
    }
    if args.copy_probe:
        out["copy_probe"] = copy_probe()
    if not args.no_extras:
        pv, psample = plain_gpu(cfg)
        out["plain_gpu"] = {"value": pv, "unit": "sarcomere-steps/s", "sample": psample}
        out["batched_over_plain"] = value / pv
        out["force_rel_err_vs_fp64"] = force_error_sample(cfg)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
