"""Philox4x32-10 counter-based generator and the stream layout (oracle side).

Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11 (Random123).  Pinned by the Random123 known-answer vectors
in tests/test_oracle_philox.py.

Stream layout shared with the HIP kernels (the reference draws from
unseeded generators instead, src/PETITE/shower.py:30, moliere.py:3):
  key     = (seed_lo, seed_hi)
  counter = (sid_lo, sid_hi, index, purpose)
where sid is a particle's 64-bit stream id.  Purposes are the PURPOSE_*
constants below.  Uniforms come from the top 23 bits:
u = ((x >> 9) + 0.5) * 2^-23, in the open interval (0, 1), exactly
representable in float32 and float64.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

PURPOSE_STEP = 0       # substep k: x hard test, y -> U(6,20), z MCS radius, w MCS azimuth (sign: bit 0 of w)
PURPOSE_CLOSE = 2      # closing increment: x fraction, y MCS radius, z MCS azimuth (sign: bit 0 of z)
PURPOSE_INTERACT = 3   # index 0: u0 process choice, u1 kinematic azimuth, u2/u3 extra
PURPOSE_VEGAS_A = 4    # candidate j: y0..y3 (ndim<=3: u3 is the accept uniform; ndim 4: low bits, see vegas)
PURPOSE_CHILD = 6      # index k: child stream id (words 0,1)
PURPOSE_DARK = 7       # index = dark process code: dark stream id (words 0,1)
PURPOSE_BOUND_E = 8    # index j: bound-electron momentum accept/reject pairs
PURPOSE_MAP_TRAIN = 9  # map training sample j of (process, energy): y0..y3


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32; key: (..., 2) uint32 -> (..., 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0 = np.broadcast_to(k[..., 0], c[..., 0].shape).astype(np.uint64)
    k1 = np.broadcast_to(k[..., 1], c[..., 0].shape).astype(np.uint64)
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        n0 = hi1 ^ c1 ^ k0
        n2 = hi0 ^ c3 ^ k1
        c0, c1, c2, c3 = n0, lo1, n2, lo0
        k0 = (k0 + np.uint64(W0)) & MASK32
        k1 = (k1 + np.uint64(W1)) & MASK32
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def to_uniform(x):
    """uint32 -> float64 uniform in (0,1), exact in float32 as well."""
    x = np.asarray(x, dtype=np.uint32)
    return ((x >> np.uint32(9)).astype(np.float64) + 0.5) * (2.0 ** -23)


def words(seed, sid, index, purpose):
    """The four raw 32-bit words of one Philox block."""
    sid = np.asarray(sid, dtype=np.uint64)
    index = np.asarray(index, dtype=np.uint64)
    shape = np.broadcast(sid, index).shape
    ctr = np.empty(shape + (4,), dtype=np.uint32)
    ctr[..., 0] = (np.broadcast_to(sid, shape) & MASK32).astype(np.uint32)
    ctr[..., 1] = (np.broadcast_to(sid, shape) >> np.uint64(32)).astype(np.uint32)
    ctr[..., 2] = (np.broadcast_to(index, shape) & MASK32).astype(np.uint32)
    ctr[..., 3] = np.uint32(purpose)
    key = np.array([int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    return philox4x32_10(ctr, key)


def block(seed, sid, index, purpose):
    """Four uniforms of one Philox block for a (scalar or array) stream."""
    sid = np.asarray(sid, dtype=np.uint64)
    index = np.asarray(index, dtype=np.uint64)
    shape = np.broadcast(sid, index).shape
    ctr = np.empty(shape + (4,), dtype=np.uint32)
    ctr[..., 0] = (np.broadcast_to(sid, shape) & MASK32).astype(np.uint32)
    ctr[..., 1] = (np.broadcast_to(sid, shape) >> np.uint64(32)).astype(np.uint32)
    ctr[..., 2] = (np.broadcast_to(index, shape) & MASK32).astype(np.uint32)
    ctr[..., 3] = np.uint32(purpose)
    key = np.array([int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    return to_uniform(philox4x32_10(ctr, key))


def derive_sid(seed, sid, index, purpose):
    """64-bit stream id derived from a parent stream (children, dark streams)."""
    sid = np.asarray(sid, dtype=np.uint64)
    index = np.asarray(index, dtype=np.uint64)
    shape = np.broadcast(sid, index).shape
    ctr = np.empty(shape + (4,), dtype=np.uint32)
    ctr[..., 0] = (np.broadcast_to(sid, shape) & MASK32).astype(np.uint32)
    ctr[..., 1] = (np.broadcast_to(sid, shape) >> np.uint64(32)).astype(np.uint32)
    ctr[..., 2] = (np.broadcast_to(index, shape) & MASK32).astype(np.uint32)
    ctr[..., 3] = np.uint32(purpose)
    key = np.array([int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    out = philox4x32_10(ctr, key)
    return out[..., 0].astype(np.uint64) | (out[..., 1].astype(np.uint64) << np.uint64(32))
