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
