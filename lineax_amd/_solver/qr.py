"""QR solver (lineax/_solver/qr.py) on the blocked Householder kernels."""

from __future__ import annotations

from .. import _kernels as K
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .misc import pack_structures, ravel_vector, transpose_packed_structures, unravel_solution


class QR(AbstractLinearSolver):
    """QR solver; handles non-square, full-rank operators.

    ``init`` factors the matrix, or its transpose when it is wide
    (lineax/_solver/qr.py:55-65), with the blocked Householder kernels.
    ``compute`` returns the least-squares solution R^{-1} (Q^T b)[:n] for a
    tall factor and the minimum-norm solution Q [R^{-T} b; 0] for a
    transposed one (qr.py:67-95); ``transpose`` flips which of the two is
    taken (qr.py:97-108).  State: ``((a, tau, tblk), transposed,
    packed_structures)`` where ``a`` holds R and the Householder vectors,
    ``tau`` the reflector scalars and ``tblk`` the compact-WY T blocks.
    """

    def init(self, operator, options):
        matrix = operator.as_matrix()
        m, n = matrix.shape[-2:]
        transpose = n > m
        if transpose:
            matrix = matrix.mT
        a, tau, tblk = K.qr_factor(matrix)
        return (a, tau, tblk), transpose, pack_structures(operator)

    def compute(self, state, vector, options):
        (a, tau, tblk), transpose, packed = state
        flat, batch = ravel_vector(vector, packed)
        x = K.qr_solve(a, tblk, flat.to(a.dtype), transpose)
        return unravel_solution(x, packed, tuple(x.shape[:-1])), RESULTS.successful, {}

    def transpose(self, state, options):
        factors, transpose, packed = state
        return (factors, not transpose, transpose_packed_structures(packed)), {}

    def conj(self, state, options):
        (a, tau, tblk), transpose, packed = state
        return ((a.conj(), tau.conj(), tblk.conj()), transpose, packed), {}

    def assume_full_rank(self):
        return True
