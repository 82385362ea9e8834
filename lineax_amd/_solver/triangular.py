"""Triangular solver (lineax/_solver/triangular.py) on the trsv kernel."""

from __future__ import annotations

from .. import _kernels as K
from .._operator import has_unit_diagonal, is_lower_triangular, is_upper_triangular
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .misc import pack_structures, ravel_vector, transpose_packed_structures, unravel_solution


class Triangular(AbstractLinearSolver):
    """Solves with a lower or upper triangular operator.

    Validation and state follow lineax/_solver/triangular.py:49-66; the solve
    (triangular.py:68-85) is the blocked trsv kernel; ``transpose`` flips the
    ``trans`` flag (triangular.py:87-99).  State: ``(matrix, lower,
    unit_diagonal, packed_structures, transposed)``.
    """

    def init(self, operator, options):
        if operator.in_size() != operator.out_size():
            raise ValueError("`Triangular` may only be used for linear solves with square matrices")
        if not (is_lower_triangular(operator) or is_upper_triangular(operator)):
            raise ValueError("`Triangular` may only be used for linear solves with triangular matrices")
        return (operator.as_matrix(), is_lower_triangular(operator), has_unit_diagonal(operator),
                pack_structures(operator), False)

    def compute(self, state, vector, options):
        matrix, lower, unit_diagonal, packed, transpose = state
        flat, batch = ravel_vector(vector, packed)
        x = K.tri_solve(matrix, flat.to(matrix.dtype), lower=lower, trans=transpose, unit_diag=unit_diagonal)
        return unravel_solution(x, packed, tuple(x.shape[:-1])), RESULTS.successful, {}

    def transpose(self, state, options):
        matrix, lower, unit_diagonal, packed, transpose = state
        return (matrix, lower, unit_diagonal, transpose_packed_structures(packed), not transpose), {}

    def conj(self, state, options):
        matrix, lower, unit_diagonal, packed, transpose = state
        return (matrix.conj(), lower, unit_diagonal, packed, transpose), {}

    def assume_full_rank(self):
        return True
