"""LU solver (lineax/_solver/lu.py) on the getrf/getrs kernels."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._operator import is_diagonal
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .misc import pack_structures, ravel_vector, transpose_packed_structures, unravel_solution


class LU(AbstractLinearSolver):
    """LU solver for square nonsingular operators.

    ``init`` factors with partial pivoting (lineax/_solver/lu.py:43-54, the
    getrf kernel); ``compute`` solves with ``trans`` taken from the state
    (lu.py:56-66), so ``transpose`` only flips a flag (lu.py:68-81).
    State: ``((lu, piv), packed_structures, transposed)``.
    """

    def init(self, operator, options):
        if operator.in_size() != operator.out_size():
            raise ValueError("`LU` may only be used for linear solves with square matrices")
        packed = pack_structures(operator)
        matrix = operator.as_matrix()
        if is_diagonal(operator):
            # lineax/_solver/lu.py:50-51: the diagonal is its own LU factor.
            n = operator.in_size()
            lu = (matrix, torch.arange(n, dtype=torch.int32, device=matrix.device).expand(*matrix.shape[:-1]))
        else:
            lu_mat, piv, _info = K.lu_factor(matrix)
            lu = (lu_mat, piv)
        return lu, packed, False

    def compute(self, state, vector, options):
        (lu, piv), packed, transpose = state
        flat, batch = ravel_vector(vector, packed)
        x = K.lu_solve(lu, piv, flat.to(lu.dtype), trans=1 if transpose else 0)
        batch = tuple(x.shape[:-1])
        return unravel_solution(x, packed, batch), RESULTS.successful, {}

    def transpose(self, state, options):
        lu_and_piv, packed, transpose = state
        return (lu_and_piv, transpose_packed_structures(packed), not transpose), {}

    def conj(self, state, options):
        # As lineax/_solver/lu.py:83-95 (which also flips the transpose flag).
        (lu, piv), packed, transpose = state
        return ((lu.conj(), piv), packed, not transpose), {}

    def assume_full_rank(self):
        return True
