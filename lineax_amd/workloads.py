"""Seeded synthetic workloads of the benchmark configurations.

All inputs come from ``numpy.random.default_rng(seed)``; float32 data are
the exact values the float64 checks upcast.

  - ``small_lu_batch``: 1024 systems A = I + G/sqrt(32), G and b iid N(0, 1),
    n = 32 (seed 0).
  - ``uneven_spd_batch``: 20,000 SPD systems, n ~ U{8..256},
    A = M M^T / n + 0.1 I built in float64 then rounded to float32,
    b iid N(0, 1) (seed 1).
"""

from __future__ import annotations

import dataclasses
import hashlib

import numpy as np


def small_lu_batch(batch: int = 1024, n: int = 32, seed: int = 0):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((batch, n, n))
    a = (np.eye(n) + g / np.sqrt(n)).astype(np.float32)
    b = rng.standard_normal((batch, n)).astype(np.float32)
    return a, b


@dataclasses.dataclass
class UnevenHost:
    """An uneven batch packed back to back (row-major matrices)."""

    n: np.ndarray  # int32 [count]
    a_offset: np.ndarray  # int64 [count]
    a: np.ndarray  # float32 [sum n^2]
    b_offset: np.ndarray  # int64 [count]
    b: np.ndarray  # float32 [sum n]

    @property
    def count(self) -> int:
        return int(self.n.shape[0])

    def matrix(self, i: int) -> np.ndarray:
        n = int(self.n[i])
        o = int(self.a_offset[i])
        return self.a[o: o + n * n].reshape(n, n)

    def rhs(self, i: int) -> np.ndarray:
        n = int(self.n[i])
        o = int(self.b_offset[i])
        return self.b[o: o + n]

    def digest(self) -> str:
        h = hashlib.sha256()
        for arr in (self.n, self.a, self.b):
            h.update(np.ascontiguousarray(arr).tobytes())
        return h.hexdigest()[:16]

    def subset(self, idx) -> "UnevenHost":
        idx = np.asarray(idx)
        ns = self.n[idx]
        a_parts = [self.a[int(self.a_offset[i]): int(self.a_offset[i]) + int(self.n[i]) ** 2] for i in idx]
        b_parts = [self.b[int(self.b_offset[i]): int(self.b_offset[i]) + int(self.n[i])] for i in idx]
        return pack_uneven(ns, a_parts, b_parts)


def pack_uneven(ns, mats, rhss) -> UnevenHost:
    ns = np.asarray(ns, np.int32)
    sq = ns.astype(np.int64) ** 2
    a_off = np.zeros(len(ns), np.int64)
    b_off = np.zeros(len(ns), np.int64)
    if len(ns):
        a_off[1:] = np.cumsum(sq)[:-1]
        b_off[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    a = np.concatenate([np.asarray(m, np.float32).reshape(-1) for m in mats]) if len(ns) else np.zeros(0, np.float32)
    b = np.concatenate([np.asarray(v, np.float32).reshape(-1) for v in rhss]) if len(ns) else np.zeros(0, np.float32)
    return UnevenHost(ns, a_off, a, b_off, b)


def uneven_spd_batch(count: int = 20000, n_min: int = 8, n_max: int = 256, seed: int = 1) -> UnevenHost:
    """SPD systems A = M M^T / n + 0.1 I (float64 build, float32 storage).

    Generated bucket by bucket (all systems of one order together), so the
    random stream is: the orders first, then per bucket in increasing n the
    M matrices, then all right-hand sides in system order.
    """
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_min, n_max + 1, size=count).astype(np.int32)
    mats: list = [None] * count
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        m = rng.standard_normal((len(idx), n, n))
        a = m @ m.transpose(0, 2, 1) / n + 0.1 * np.eye(n)
        a32 = a.astype(np.float32)
        for k, i in enumerate(idx):
            mats[i] = a32[k]
    rhss = [rng.standard_normal(int(n)).astype(np.float32) for n in ns]
    return pack_uneven(ns, mats, rhss)
