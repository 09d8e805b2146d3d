"""Philox4x32-10 counter-based generator (NumPy), test infrastructure only.

Published algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
as easy as 1, 2, 3", SC'11 -- 10 rounds of the 4x32 Philox bijection with
multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
0xBB67AE85.  Known-answer vectors from the paper's reference implementation
are checked in tests/test_oracle_rng.py.

The device kernels draw their uniforms from this generator; the oracle
regenerates the same draws to replay a run with identical uniforms (the
reference's own seam for deterministic testing: ``random_values=`` in
multifil_jax/kernels/transitions.py:521,1086).

Draw layout (shared with the device code):
    key     = (seed & 0xffffffff, seed >> 32)
    counter = (block, step, purpose, 0)   purpose 0 = TM sites, 1 = XBs
    uniform[4*block + k] = (word_k >> 9) * 2**-23    (a float32 in [0, 1))
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

PURPOSE_TM = 0
PURPOSE_XB = 1


def philox4x32_10(ctr, key):
    """ctr: (4, n) uint32; key: (2,) or (2, n) uint32 -> (4, n) uint32."""
    c0, c1, c2, c3 = [np.asarray(c, dtype=np.uint32).copy() for c in ctr]
    k0 = np.asarray(key[0], dtype=np.uint32).copy()
    k1 = np.asarray(key[1], dtype=np.uint32).copy()
    with np.errstate(over='ignore'):
        for r in range(10):
            p0 = c0.astype(np.uint64) * M0
            p1 = c2.astype(np.uint64) * M1
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            if r < 9:
                k0 = (k0 + W0).astype(np.uint32)
                k1 = (k1 + W1).astype(np.uint32)
    return np.stack([c0, c1, c2, c3])


def uniforms(seed: int, step: int, purpose: int, n: int) -> np.ndarray:
    """The n float32 uniforms an instance draws for (step, purpose)."""
    nblk = (n + 3) // 4
    blk = np.arange(nblk, dtype=np.uint32)
    ctr = np.stack([blk,
                    np.full(nblk, step & 0xFFFFFFFF, np.uint32),
                    np.full(nblk, purpose, np.uint32),
                    np.zeros(nblk, np.uint32)])
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    words = philox4x32_10(ctr, key)              # (4, nblk)
    bits = words.T.reshape(-1)[:n]               # element 4*blk + k = word k
    return ((bits >> np.uint32(9)).astype(np.float32) * np.float32(2.0 ** -23)).astype(np.float32)
