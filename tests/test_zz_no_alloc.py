"""mf_run_noalloc (the allocation-free entry point) against mf_run, on the
GPU, through the C ABI.

The 6x6 lattice runs the device-workspace kernels, the only ones with a
scratch buffer: mf_run takes a workspace the batch allocates, mf_run_noalloc
the library's static array.  Two calls in a row, the second with fewer
instances (so it reuses the array the first left dirty); each must equal
mf_run bit for bit (same kernel, same inputs): forces, all 52 metrics and
the final states.  A 2x2 launch (no scratch) must match too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(topo, d, B, T, seed0, noalloc):
    from multifil_jax_amd.engine import LatticePack, Batch, SolverConfig
    drv = np.zeros((B, T, 3), np.float32)
    drv[..., 0] = np.linspace(950.0, 1050.0, B, dtype=np.float32)[:, None]
    drv[..., 1] = np.linspace(4.5, 5.5, B, dtype=np.float32)[:, None]
    drv[..., 2] = 14.0
    bt = Batch(LatticePack([topo]), np.zeros(B, np.int32), d.rows(B), drv,
               seeds=np.arange(seed0, seed0 + B, dtype=np.uint64))
    bt.run(T, cfg=SolverConfig(mode="refined"), noalloc=noalloc)
    return bt


def _same(a, b):
    assert np.array_equal(a.force(), b.force())
    assert np.array_equal(a.metrics(), b.metrics(), equal_nan=True)
    for i in range(a.B):
        sa, sb = a.get_state(i), b.get_state(i)
        for k in sa:
            assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), k


def test_noalloc_matches_mf_run(skeletal):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    topo = SarcTopology.create(6, 6, s, d)
    for B, seed0 in ((3, 40), (2, 50)):
        print(f"noalloc 6x6: B={B}, 6 steps", flush=True)
        ref = _pair(topo, d, B, 6, seed0, noalloc=False)
        assert ref.d_ws is not None            # the workspace kernels ran
        got = _pair(topo, d, B, 6, seed0, noalloc=True)
        assert got.d_ws is None                # ... with no workspace from the batch
        _same(ref, got)
    print("noalloc 2x2: B=4, 8 steps", flush=True)
    t22 = SarcTopology.create(2, 2, s, d)
    _same(_pair(t22, d, 4, 8, 60, noalloc=False), _pair(t22, d, 4, 8, 60, noalloc=True))
