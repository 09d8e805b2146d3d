"""Write the committed fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_fixtures.py``).

* ``conditions.json``: the condition tables of the measured workloads, drawn
  with numpy.random.default_rng(20261203):
  - batched sweep: 256 x (pCa ~ U[4, 7], z_line ~ U[850, 1150] nm);
  - uneven batch: 96 instances of each lattice (2x2, 3x3, 4x4, 6x6),
    lengths ~ U{200..1000} steps, pCa ~ U[4.5, 6.5];
  - sinusoid: 64 frequencies logspace(0, 3) Hz.
* ``oracle_2x2_pCa45.npz``: 8 steps of the FP64 oracle (exact solves) on the
  2x2 lattice at pCa 4.5, z_line 1000 nm, seed 5: the 52 metrics per step.
  A regression pin of the oracle itself, not a reference output (the
  reference has no recorded trajectories).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED = 20261203


def conditions():
    rng = np.random.default_rng(SEED)
    pCa = rng.uniform(4.0, 7.0, 256).astype(np.float32)
    z = rng.uniform(850.0, 1150.0, 256).astype(np.float32)
    rng = np.random.default_rng(SEED)
    lat = np.repeat(np.arange(4), 96)
    lengths = rng.integers(200, 1001, 384)
    pCa_u = rng.uniform(4.5, 6.5, 384).astype(np.float32)
    return {
        "seed": SEED,
        "sweep": {"pCa": pCa.tolist(), "z_line": z.tolist()},
        "uneven": {"lattices": [2, 3, 4, 6], "lattice_index": lat.tolist(), "lengths": lengths.tolist(),
                   "pCa": pCa_u.tolist()},
        "sinusoid": {"freq_hz": np.logspace(0, 3, 64).tolist()},
    }


def oracle_trajectory():
    from oracle import mf_oracle as mo
    from multifil_jax_amd.params import get_skeletal_params, DYNAMIC_FIELDS
    from multifil_jax_amd.topology import SarcTopology
    s, d = get_skeletal_params()
    topo = SarcTopology.create(2, 2, s, d)
    p = {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}
    T = 8
    _, out = mo.run(topo, p, np.full(T, 1000.0), np.full(T, 4.5), np.full(T, 14.0), 1.0, 5,
                    mo.SolverConfig(mode="exact"))
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "conditions.json"), "w") as f:
        json.dump(conditions(), f)
    np.savez(os.path.join(HERE, "oracle_2x2_pCa45.npz"), metrics=oracle_trajectory(),
             seed=np.int64(5), z_line=np.float64(1000.0), pCa=np.float64(4.5))
    print("fixtures written to", HERE)
