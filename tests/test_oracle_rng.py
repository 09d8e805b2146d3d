"""Philox4x32-10 known-answer vectors (Salmon et al., SC'11 reference
implementation, Random123 kat_vectors) and the uniform mapping."""
import numpy as np

from oracle.philox import philox4x32_10, uniforms, PURPOSE_TM, PURPOSE_XB

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_philox_known_answers():
    for ctr, key, want in KAT:
        out = philox4x32_10(np.array(ctr, np.uint32).reshape(4, 1), np.array(key, np.uint32))
        assert [int(w[0]) for w in out] == want


def test_uniform_range_and_layout():
    u = uniforms(7, 3, PURPOSE_XB, 1001)
    assert u.dtype == np.float32 and u.shape == (1001,)
    assert np.all(u >= 0) and np.all(u < 1)
    # 23-bit grid: u * 2^23 is an integer
    assert np.all(np.floor(u.astype(np.float64) * 2 ** 23) == u.astype(np.float64) * 2 ** 23)
    # element 4*blk + k is word k of block blk
    w = philox4x32_10(np.array([[5], [3], [PURPOSE_XB], [0]], np.uint32), np.array([7, 0], np.uint32))
    assert np.allclose(u[20:24], (w[:, 0] >> 9).astype(np.float64) * 2.0 ** -23)


def test_streams_differ():
    a = uniforms(0, 0, PURPOSE_TM, 64)
    b = uniforms(0, 0, PURPOSE_XB, 64)
    c = uniforms(1, 0, PURPOSE_TM, 64)
    d = uniforms(0, 1, PURPOSE_TM, 64)
    assert not np.array_equal(a, b) and not np.array_equal(a, c) and not np.array_equal(a, d)
    # moments of a larger sample
    u = uniforms(3, 0, PURPOSE_TM, 200000).astype(np.float64)
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
