"""Diagonal solver (lineax/_solver/diagonal.py)."""

from __future__ import annotations

import torch

from .._misc import resolve_rcond
from .._operator import diagonal, has_unit_diagonal, is_diagonal
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .misc import pack_structures, ravel_vector, transpose_packed_structures, unravel_solution


class Diagonal(AbstractLinearSolver):
    """x = b / diag(A); with ``well_posed=False`` tiny entries give zeros (pseudoinverse)."""

    def __init__(self, well_posed: bool = False, rcond: float | None = None):
        self.well_posed = well_posed
        self.rcond = rcond

    def init(self, operator, options):
        if operator.in_size() != operator.out_size():
            raise ValueError("`Diagonal` may only be used for linear solves with square matrices")
        if not is_diagonal(operator):
            raise ValueError("`Diagonal` may only be used for linear solves with diagonal matrices")
        packed = pack_structures(operator)
        if has_unit_diagonal(operator):
            return None, packed
        return diagonal(operator), packed

    def compute(self, state, vector, options):
        diag, packed = state
        flat, batch = ravel_vector(vector, packed)
        if diag is None:
            x = flat
        else:
            d = diag.to(flat.device)
            if not self.well_posed:
                size = d.shape[-1]
                rcond = resolve_rcond(self.rcond, size, size, d.dtype)
                ad = torch.abs(d)
                cutoff = rcond * ad.amax(dim=-1, keepdim=True)
                d = torch.where(ad > cutoff, d, torch.full_like(d, float("inf")))
            x = flat / d
        return unravel_solution(x, packed, tuple(x.shape[:-1])), RESULTS.successful, {}

    def transpose(self, state, options):
        diag, packed = state
        return (diag, transpose_packed_structures(packed)), {}

    def conj(self, state, options):
        diag, packed = state
        return (None if diag is None else diag.conj(), packed), {}

    def assume_full_rank(self):
        return self.well_posed
