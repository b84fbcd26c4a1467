"""Seeded probe generator: xoshiro256++ (Blackman & Vigna, 2019) seeded by
splitmix64, vectorised over independent lanes.

This is the build's own documented generator.  It is *not* Julia's
``Xoshiro(seed)`` stream (Julia seeds its xoshiro256++ state differently and
its bulk ``rand!`` interleaves several generators), so probes made here are
not the reference's probes; parity runs pass identical probe arrays to both
sides instead.

Lane layout: lane j's state is seeded from splitmix64 outputs 4j..4j+3 of the
stream started at ``seed``.  ``rand(n)`` returns value i from lane i % lanes,
step i // lanes; arrays are filled in column-major (Julia) order.
"""
from __future__ import annotations

import numpy as np

_U64 = np.uint64


def splitmix64_stream(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of splitmix64 started at ``seed``."""
    out = np.empty(count, dtype=np.uint64)
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out[i] = z ^ (z >> 31)
    return out


def _rotl(x: np.ndarray, k: int) -> np.ndarray:
    return (x << _U64(k)) | (x >> _U64(64 - k))


class Xoshiro256pp:
    """Lane-parallel xoshiro256++."""

    def __init__(self, seed: int = 42, lanes: int = 1024, state: np.ndarray | None = None):
        if state is not None:
            st = np.asarray(state, dtype=np.uint64).reshape(4, -1)
        else:
            st = splitmix64_stream(seed, 4 * lanes).reshape(lanes, 4).T.copy()
        self.s = [st[0].copy(), st[1].copy(), st[2].copy(), st[3].copy()]
        self.lanes = self.s[0].shape[0]

    def next_u64(self) -> np.ndarray:
        """One output per lane."""
        s0, s1, s2, s3 = self.s
        with np.errstate(over="ignore"):
            result = _rotl(s0 + s3, 23) + s0
        t = s1 << _U64(17)
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        self.s[3] = _rotl(s3, 45)
        return result

    def random_u64(self, n: int) -> np.ndarray:
        steps = -(-n // self.lanes)
        out = np.empty((steps, self.lanes), dtype=np.uint64)
        for k in range(steps):
            out[k] = self.next_u64()
        return out.reshape(-1)[:n]

    def rand(self, *shape) -> np.ndarray:
        """Uniform doubles in [0, 1) with 53 random bits, column-major fill."""
        n = int(np.prod(shape)) if shape else 1
        u = self.random_u64(n)
        v = (u >> _U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        if not shape:
            return float(v[0])
        return v.reshape(shape[::-1]).T if len(shape) > 1 else v


def default_rng(seed: int = 42) -> Xoshiro256pp:
    return Xoshiro256pp(seed)
