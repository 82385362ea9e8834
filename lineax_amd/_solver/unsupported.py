"""Solvers of lineax that this package does not provide.

They exist so that ``AutoLinearSolver`` dispatch and imports behave as in
lineax; using one raises ``NotImplementedError`` naming the reference file.
"""

from __future__ import annotations

from .._solve import AbstractLinearSolver


class _Unsupported(AbstractLinearSolver):
    _reference = ""

    def __init__(self, *args, **kwargs):
        self.args = args
        self.kwargs = kwargs

    def _fail(self):
        raise NotImplementedError(
            f"{type(self).__name__} ({self._reference}) is not provided by lineax_amd; "
            "use LU, Cholesky, QR, CG or GMRES"
        )

    def init(self, operator, options):
        self._fail()

    def compute(self, state, vector, options):
        self._fail()

    def transpose(self, state, options):
        self._fail()

    def conj(self, state, options):
        self._fail()

    def assume_full_rank(self):
        return False


class Tridiagonal(_Unsupported):
    _reference = "lineax/_solver/tridiagonal.py"

    def assume_full_rank(self):
        return True
