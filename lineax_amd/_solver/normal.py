"""Normal-equations wrapper (lineax/_solver/normal.py) around an inner solver
of positive (semi)definite systems."""

from __future__ import annotations

from copy import copy

from .._operator import TaggedLinearOperator, conj, linearise, materialise
from .._solve import AbstractLinearSolver
from .._tags import positive_semidefinite_tag
from .cholesky import Cholesky


def _normal_options(options: dict, tall: bool) -> dict:
    """The inner solver's options (lineax/_solver/normal.py:31-50): a
    preconditioner M becomes M M^H (tall) or M^H M (wide), and in the wide
    form the initial guess y0 becomes M^H y0."""
    inner = copy(options)
    preconditioner = options.get("preconditioner")
    y0 = options.get("y0")
    if preconditioner is not None:
        preconditioner = linearise(preconditioner)
        if tall:
            inner["preconditioner"] = TaggedLinearOperator(preconditioner @ conj(preconditioner.transpose()),
                                                           positive_semidefinite_tag)
        else:
            inner["preconditioner"] = TaggedLinearOperator(conj(preconditioner.transpose()) @ preconditioner,
                                                           positive_semidefinite_tag)
            if y0 is not None:
                inner["y0"] = conj(preconditioner.transpose()).mv(y0)
    return inner


class Normal(AbstractLinearSolver):
    """Solves A x = b through the normal equations with ``inner_solver``:
    A^H A x = A^H b when m >= n, else A A^H y = b with x = A^H y
    (lineax/_solver/normal.py:53-104).

    ``init`` materialises the operator when the inner solver is Cholesky, so
    the inner operator's matrix A^H A (or A A^H) is one device GEMM and the
    factorisation runs on the Cholesky kernels; for other inner solvers
    (e.g. CG) the operator is only linearised and the products stay
    matrix-free (normal.py:105-120).  ``compute`` applies A^H before (tall)
    or after (wide) the inner solve (normal.py:122-140); ``transpose`` and
    ``conj`` conjugate the inner state and swap or keep the tall flag
    (normal.py:142-179).  State: ``(inner_state, tall,
    operator_conj_transpose, inner_options)``.
    """

    def __init__(self, inner_solver: AbstractLinearSolver):
        self.inner_solver = inner_solver

    def init(self, operator, options):
        tall = operator.out_size() >= operator.in_size()
        lin_op = materialise(operator) if isinstance(self.inner_solver, Cholesky) else linearise(operator)
        if tall:
            inner_operator = conj(lin_op.transpose()) @ lin_op
        else:
            inner_operator = lin_op @ conj(lin_op.transpose())
        inner_operator = TaggedLinearOperator(inner_operator, positive_semidefinite_tag)
        inner_options = _normal_options(options, tall)
        inner_state = self.inner_solver.init(inner_operator, inner_options)
        return inner_state, tall, conj(lin_op.transpose()), inner_options

    def compute(self, state, vector, options):
        inner_state, tall, operator_conj_transpose, inner_options = state
        if tall:
            vector = operator_conj_transpose.mv(vector)
        solution, result, stats = self.inner_solver.compute(inner_state, vector, inner_options)
        if not tall:
            solution = operator_conj_transpose.mv(solution)
        return solution, result, stats

    def transpose(self, state, options):
        inner_state, tall, operator_conj_transpose, inner_options = state
        inner_state, inner_options = self.inner_solver.conj(inner_state, inner_options)
        return (inner_state, not tall, operator_conj_transpose.transpose(), inner_options), options

    def conj(self, state, options):
        inner_state, tall, operator_conj_transpose, inner_options = state
        inner_state, inner_options = self.inner_solver.conj(inner_state, inner_options)
        return (inner_state, tall, conj(operator_conj_transpose), inner_options), options

    def assume_full_rank(self):
        return self.inner_solver.assume_full_rank()
