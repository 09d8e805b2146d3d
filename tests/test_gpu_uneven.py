"""The uneven workload at full size on the GPU: 384 instances, 96 each of
2x2 / 3x3 / 4x4 / 6x6, lengths U{200..1000} steps, pCa U[4.5, 6.5],
dynamic lattice spacing (K_lat 5 pN/nm per thick filament, nu 0), refined
solver, autotuned, in the benchmark's two concurrent launches.

Size-independent properties checked over every instance and step:
* the refined residual is <= 1e-5 pN at every step of every instance;
* XB state counts sum to the lattice's X heads and TM counts to its S
  sites; n_bound equals the count of states 1-3;
* the solved spacing stays finite and inside (5, 25) nm;
* rows past an instance's own length stay exactly zero (metrics and the
  FP64 force);
* the 6x6 launch ran the device-workspace variant.
Per-step parity of these lattices against the oracle is covered by the
dynamic-spacing tests at 2x2, 3x3, 4x4 and 6x6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_uneven_workload_full_size():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    import bench
    from multifil_jax_amd.engine import SolverConfig
    print("uneven workload: building and autotuning two launches", flush=True)
    batches, T, total = bench.uneven_batches(SolverConfig(mode="refined"))
    assert sum(bt.B for bt in batches) == 384 and T <= 1000
    el = bench.run_uneven(batches, T)
    print(f"uneven workload: {total} sarcomere-steps in {el:.2f} s", flush=True)
    assert batches[0].d_ws is not None       # 6x6: workspace variant
    n_lat = {}
    for bt in batches:
        m, f = bt.metrics(), bt.force()
        for b in range(bt.B):
            t = bt.pack.topos[bt.lattice_ids[b]]
            n_lat[t.n_thick] = n_lat.get(t.n_thick, 0) + 1
            L = int(bt.lengths[b])
            X, S = t.total_xbs, t.n_thin * t.n_sites
            mb = m[b, :L]
            assert np.all(mb[:, 1] <= 1e-5), (b, float(mb[:, 1].max()))
            assert np.all(mb[:, 6:12].sum(axis=1) == X)
            assert np.all(mb[:, 19:23].sum(axis=1) == S)
            assert np.all(mb[:, 5] == mb[:, 7:10].sum(axis=1))
            assert np.all(np.isfinite(mb[:, 4]) & (mb[:, 4] > 5.0) & (mb[:, 4] < 25.0))
            assert np.all(np.isfinite(f[b, :L]))
            assert np.all(m[b, L:] == 0.0) and np.all(f[b, L:] == 0.0)
    assert sorted(n_lat.values()) == [96, 96, 96, 96]
