"""Host-side logic of the batch engine that needs no GPU."""
import numpy as np

from multifil_jax_amd.engine import Batch, default_block


class _Drivers:
    """The two attributes Batch.segments reads."""
    def __init__(self, drv):
        self.drivers_np = drv
        self._changes = None


def test_segments_split_at_pca_and_spacing_changes():
    drv = np.zeros((3, 12, 3), np.float32)
    drv[..., 1], drv[..., 2] = 4.5, 14.0
    drv[1, 5:, 1] = 5.0          # pCa change at step 5 in one instance
    drv[2, 9:, 2] = 13.0         # spacing change at step 9 in another
    drv[0, :, 0] = np.arange(12)  # Z-line changes never split
    b = _Drivers(drv)
    assert Batch.segments(b, 0, 12) == [(0, 5), (5, 4), (9, 3)]
    assert Batch.segments(b, 3, 5) == [(3, 2), (5, 3)]
    assert Batch.segments(b, 5, 4) == [(5, 4)]
    assert Batch.segments(b, 0, 0) == []


def test_default_block(topo22, topo44):
    assert default_block([topo22]) == 768          # 12 chains: one wave each
    assert default_block([topo44]) == 1024
    assert default_block([topo22, topo44]) == 1024


def test_nan_drivers_fall_back_to_the_parameter_row(skeletal):
    """core/state.py:331-343: a NaN driver means "use the constant"."""
    from multifil_jax_amd.engine import resolve_drivers
    _, d = skeletal
    rows = d.rows(2)
    drv = np.zeros((2, 4, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = 1000.0, 5.0, 13.0
    assert resolve_drivers(drv, rows) is drv
    holes = drv.copy()
    holes[0, 1, 0] = holes[1, 2, 1] = holes[1, 3, 2] = np.nan
    out = resolve_drivers(holes, rows)
    want = drv.copy()
    want[0, 1, 0] = np.float32(d.z_line)
    want[1, 2, 1] = np.float32(d.pCa)
    want[1, 3, 2] = np.float32(d.lattice_spacing)
    assert out.dtype == np.float32 and np.array_equal(out, want)


def _host_batch(topo, d, pops=None, B=2):
    drv = np.zeros((B, 5, 3), np.float32)
    drv[..., 0], drv[..., 1], drv[..., 2] = 900.0, 4.5, 14.0
    from multifil_jax_amd.engine import LatticePack
    return Batch(LatticePack([topo]), np.zeros(B, np.int32), d.rows(B), drv,
                 seeds=np.arange(B, dtype=np.uint64), device="cpu", pops=pops)


def test_autotune_blocks_with_subpopulations(topo22, skeletal):
    """The autotuner's probe launch names the population buffers, so every
    block size stays feasible (host-side variant choice, no GPU call)."""
    from multifil_jax_amd.engine import Populations
    from multifil_jax_amd.autotune import feasible_blocks
    _, d = skeletal
    B = 2
    P = np.repeat(d.rows(B)[:, None, :], 2, axis=1).copy()
    X, S = topo22.total_xbs, topo22.n_thin * topo22.n_sites
    for has_xb, has_tm in ((True, False), (False, True), (True, True)):
        pops = Populations("explicit", P, xb_labels=[np.zeros(X, np.uint8)] * B,
                           tm_labels=[np.zeros(S, np.uint8)] * B, has_xb=has_xb, has_tm=has_tm)
        assert feasible_blocks(_host_batch(topo22, d, pops), "refined") == [512, 768, 1024]
    mf = Populations("mean_field", P, fractions=np.full((B, 2), 0.5, np.float32))
    assert feasible_blocks(_host_batch(topo22, d, mf), "refined") == [512, 768, 1024]


def test_autotune_cache_key(topo22, skeletal):
    """The cache key separates batch sizes, conditions and fixed settings."""
    from multifil_jax_amd.autotune import _key
    from multifil_jax_amd.engine import SolverConfig
    _, d = skeletal
    a, b = _host_batch(topo22, d, B=2), _host_batch(topo22, d, B=3)
    base = SolverConfig(mode="refined")
    assert _key(a, base) != _key(b, base)
    assert _key(a, base) != _key(a, SolverConfig(mode="refined", dt=0.1))
    c = _host_batch(topo22, d, B=2)
    assert _key(a, base) == _key(c, base)
    c.drivers_np = c.drivers_np.copy()
    c.drivers_np[0, 0, 1] = 5.0
    assert _key(a, base) != _key(c, base)
