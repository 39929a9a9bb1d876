"""Vectorised Philox4x32-10 (the construction of openrand::Philox,
src/PatchOperation/Random/philox.h:60-139) in numpy, for seeded synthetic inputs.

Initial conditions draw from streams of their own (STREAM_* below), never from the
integrators' streams 0x5324120 (BD) and 0x1356914 (BAOAB).
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

STREAM_POSITIONS = 0x494E4954  # "INIT"
STREAM_MOMENTA = 0x4D4F4D54    # "MOMT"
STREAM_CHAINS = 0x43484E53     # "CHNS"
STREAM_SIZES = 0x53495A45      # "SIZE"
STREAM_NOISE = 0x4E4F4953      # "NOIS"


def philox4x32_10(seed: int, ctr0, ctr1, ctr2, ctr3=0):
    """Return 4 uint32 arrays for counters (ctr0, ctr1, ctr2, ctr3) under a 64-bit key
    (key = {seed >> 32, seed & 0xffffffff})."""
    shape = np.broadcast(ctr0, ctr1, ctr2, ctr3).shape
    c = [np.broadcast_to(np.asarray(x, dtype=np.uint64) & MASK, shape).copy() for x in (ctr0, ctr1, ctr2, ctr3)]
    k0 = np.uint32((seed >> 32) & 0xFFFFFFFF)
    k1 = np.uint32(seed & 0xFFFFFFFF)
    for r in range(10):
        if r > 0:
            k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return [x.astype(np.uint32) for x in c]


def u01(x: np.ndarray) -> np.ndarray:
    """u01<float, uint32_t> (base_state.h:299-305): float(x) * 2^-32 + 2^-33."""
    f = x.astype(np.float32)
    return (f * np.float32(2.0 ** -32) + np.float32(2.0 ** -33)).astype(np.float32)


def uniform(seed: int, stream: int, n: int, draw: int = 0, offset: int = 0):
    """n draws of 4 uniforms in (0, 1], counter (offset + i, stream, draw, 0)."""
    i = np.arange(offset, offset + n, dtype=np.uint64)
    out = philox4x32_10(seed, i, stream, draw, 0)
    return np.stack([u01(o) for o in out], axis=1)


def normal(seed: int, stream: int, n: int, draw: int = 0, offset: int = 0) -> np.ndarray:
    """n x 3 standard normals via the Box-Muller mapping of BDIntegrate (BD.h:73-78),
    evaluated in float64."""
    u = uniform(seed, stream, n, draw, offset).astype(np.float64)
    r1 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r2 = np.sqrt(-2.0 * np.log(u[:, 2]))
    t1 = 2.0 * np.pi * u[:, 1]
    t2 = 2.0 * np.pi * u[:, 3]
    return np.stack([r1 * np.cos(t1), r1 * np.sin(t1), r2 * np.cos(t2)], axis=1)
