"""Cholesky solver (lineax/_solver/cholesky.py) on the potrf/potrs kernels."""

from __future__ import annotations

from .. import _kernels as K
from .._operator import is_negative_semidefinite, is_positive_semidefinite
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .._structure import batch_shape_of, ravel, unravel


class Cholesky(AbstractLinearSolver):
    """Cholesky solver for positive or negative definite operators.

    ``init`` factors A = U^T U (or -A for NSD operators,
    lineax/_solver/cholesky.py:43-63); ``compute`` solves and negates for NSD
    (cholesky.py:65-78).  A non-positive pivot yields a NaN factor and hence
    RESULTS.singular.  State: ``(U, is_nsd, structure)``; lineax keeps no
    structure because JAX flattens the whole vector, while here the
    structure separates batch dimensions from the vector's own shape.
    """

    def init(self, operator, options):
        is_nsd = is_negative_semidefinite(operator)
        if not (is_positive_semidefinite(operator) or is_nsd):
            raise ValueError(
                "`Cholesky(..., normal=False)` may only be used for positive or negative definite linear operators"
            )
        matrix = operator.as_matrix()
        m, n = matrix.shape[-2:]
        if m != n:
            raise ValueError("`Cholesky(..., normal=False)` may only be used for linear solves with square matrices")
        factor, _info = K.chol_factor(matrix, negate=is_nsd)
        return factor, is_nsd, operator.in_structure()

    def compute(self, state, vector, options):
        factor, is_nsd, structure = state
        batch = batch_shape_of(vector, structure)
        flat = ravel(vector, structure, batch).to(factor.dtype)
        x = K.chol_solve(factor, flat, negate=is_nsd)
        return unravel(x, structure, tuple(x.shape[:-1])), RESULTS.successful, {}

    def transpose(self, state, options):
        # The operator is self-adjoint (lineax/_solver/cholesky.py:80-85).
        factor, is_nsd, structure = state
        return (factor.conj(), is_nsd, structure), options

    def conj(self, state, options):
        factor, is_nsd, structure = state
        return (factor.conj(), is_nsd, structure), options

    def assume_full_rank(self):
        return True
