"""Uneven batches: many independent systems of different orders in one launch.

lineax can batch only systems of one shape (``jax.vmap``), so an uneven mix
becomes one vmap per size bucket (the "plain form" here: one FP32 launch per
bucket).  ``UnevenBatch`` keeps the whole mix resident in HBM, packed back to
back, and ``cholesky_solve`` runs it as one persistent launch per
register-tile class: workgroups pull systems from a device queue sorted
largest-first, factor in FP32 (registers and LDS) and refine the solution in
FP64 until the normwise backward error meets ``tol``.

Each system i means ``linear_solve(MatrixLinearOperator(A_i,
positive_semidefinite_tag), b_i, Cholesky())`` (lineax/_solver/cholesky.py)
for the exact float64 value of its float32 data.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import _kernels as K
from ._solution import RESULTS

MAX_ORDER = 256


@dataclasses.dataclass
class UnevenSolution:
    x: torch.Tensor  # packed solutions (float64 chosen form / float32 plain form)
    result: torch.Tensor  # int32 [count], RESULTS codes
    info: torch.Tensor  # int32 [count], 0 or first non-positive pivot (1-based)
    iters: torch.Tensor  # int32 [count], FP64 residuals formed
    eta: torch.Tensor  # float64 [count], last normwise backward error (-1 if none)

    def solution(self, batch: "UnevenBatch", i: int) -> torch.Tensor:
        o = int(batch.b_offset_host[i])
        return self.x[o: o + int(batch.n_host[i])]


class UnevenBatch:
    """Device-resident uneven batch of square systems (orders 1..256)."""

    def __init__(self, n, a_offset, a, b_offset, b, device="cuda"):
        n = np.ascontiguousarray(n, np.int32)
        if n.size and (n.min() < 1 or n.max() > MAX_ORDER):
            raise ValueError(f"orders must lie in [1, {MAX_ORDER}]")
        a_offset = np.ascontiguousarray(a_offset, np.int64)
        b_offset = np.ascontiguousarray(b_offset, np.int64)
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        if n.size:
            if np.any(a_offset < 0) or np.any(a_offset + n.astype(np.int64) ** 2 > a.size):
                raise ValueError("matrix offsets run past the packed matrix storage")
            if np.any(b_offset < 0) or np.any(b_offset + n > b.size):
                raise ValueError("vector offsets run past the packed vector storage")
        self.count = int(n.size)
        self.n_host = n
        self.b_offset_host = b_offset
        self.max_n = int(n.max()) if n.size else 1
        self.x_len = int(b.size)
        dev = torch.device(device)
        self.device = dev
        self.n = torch.from_numpy(n).to(dev)
        self.a_offset = torch.from_numpy(a_offset).to(dev)
        self.a = torch.from_numpy(a).to(dev)
        self.b_offset = torch.from_numpy(b_offset).to(dev)
        self.b = torch.from_numpy(b).to(dev)
        # Largest first for the persistent queue (ties by index: deterministic).
        desc = np.lexsort((np.arange(n.size), -n.astype(np.int64))).astype(np.int32)
        self.order_desc = torch.from_numpy(desc).to(dev)
        # Per-order buckets for the plain form.
        asc = np.lexsort((np.arange(n.size), n)).astype(np.int32)
        self.order_asc = torch.from_numpy(asc).to(dev)
        self.buckets = []
        if n.size:
            sorted_n = n[asc]
            starts = np.flatnonzero(np.r_[True, sorted_n[1:] != sorted_n[:-1]])
            ends = np.r_[starts[1:], n.size]
            self.buckets = [(int(sorted_n[s]), int(s), int(e)) for s, e in zip(starts, ends)]
        self._workspace = torch.empty(256, dtype=torch.uint8, device=dev)

    @classmethod
    def from_host(cls, host, device="cuda"):
        return cls(host.n, host.a_offset, host.a, host.b_offset, host.b, device=device)

    @property
    def algorithmic_bytes(self) -> int:
        """4 (n^2 + 2n) bytes per system: the float32 matrix, b and x."""
        n = self.n_host.astype(np.int64)
        return int(np.sum(4 * (n * n + 2 * n)))

    @property
    def algorithmic_flops(self) -> float:
        """n^3/3 + 2 n^2 flops per system (factor + two triangular solves)."""
        n = self.n_host.astype(np.float64)
        return float(np.sum(n ** 3 / 3 + 2 * n ** 2))

    def _outputs(self, x_dtype):
        dev = self.device
        return (
            torch.empty(self.x_len, dtype=x_dtype, device=dev),
            torch.empty(self.count, dtype=torch.int32, device=dev),
            torch.zeros(self.count, dtype=torch.int32, device=dev),
            torch.full((self.count,), -1.0, dtype=torch.float64, device=dev),
        )

    def cholesky_solve(self, *, form: str = "chosen", tol: float = 1e-12, max_refine: int = 10,
                       out: UnevenSolution | None = None) -> UnevenSolution:
        """Solves every system.

        ``form="chosen"``: persistent launches (one per tile class), FP32 factorisation, FP64
        refinement to ``tol``; float64 solutions.
        ``form="plain"``: one non-persistent FP32 launch per order bucket, no
        refinement; float32 solutions.
        """
        if form not in ("chosen", "plain"):
            raise ValueError("form must be 'chosen' or 'plain'")
        x_dtype = torch.float64 if form == "chosen" else torch.float32
        if out is None or out.x.dtype != x_dtype:
            x, info, iters, eta = self._outputs(x_dtype)
        else:
            x, info, iters, eta = out.x, out.info, out.iters, out.eta
        if self.count:
            if form == "chosen":
                K.chol_uneven_solve(
                    self.n, self.a_offset, self.a, self.b_offset, self.b, x, self.order_desc,
                    max_n=self.max_n, persistent=True, max_refine=max_refine, tol=tol, info=info,
                    iters=iters, eta=eta, workspace=self._workspace, sorted_order=True,
                )
            else:
                for order_n, s, e in self.buckets:
                    K.chol_uneven_solve(
                        self.n, self.a_offset, self.a, self.b_offset, self.b, x, self.order_asc[s:e],
                        max_n=order_n, persistent=False, max_refine=0, tol=0.0, info=info,
                        iters=iters, eta=eta,
                    )
        result = torch.where(info != 0, RESULTS.singular, RESULTS.successful).to(torch.int32)
        return UnevenSolution(x=x, result=result, info=info, iters=iters, eta=eta)
