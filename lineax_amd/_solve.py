"""``linear_solve``, the solver interface, and differentiation through solves.

Mirrors lineax/_solve.py:
  - ``_linear_solve_impl`` (97-129): ``solver.compute`` then the RESULTS
    post-processing (non-finite output -> singular, non-finite input ->
    nonfinite_input) on the device, then ``throw``.
  - ``linear_solve`` (656-806): validation, ``init`` on a stop-gradient
    operator unless ``state`` is given, and the solve.
  - ``AbstractLinearSolver`` (343-480) and ``AutoLinearSolver`` (518-645).
  - ``invert`` (809-871).

Differentiation (lineax's JVP rule, 151-268, and transpose rule, 286-317)
is attached to a solution x* computed without autograd:

    x = x* + Z(b - A x*)

where A and b carry the caller's autograd history and Z is a function whose
value is zero, whose forward-mode derivative is a solve with the same state
(A^+ applied to the tangent) and whose reverse-mode derivative is the solve
with ``solver.transpose(state)``.  So x' = A^+(b' - A' x*) and
ct_b = A^{+T} ct_x, with the factorisation reused exactly as lineax's rules
reuse ``state``; for non-square operators the extra pseudoinverse terms of
lineax/_solve.py:213-253 are added in the same way.  Derivatives work with
``torch.autograd`` (backward) and ``torch.autograd.forward_ad`` (JVP).
"""

from __future__ import annotations

import abc
from typing import Any

import torch

from . import _kernels as K
from ._operator import (
    AbstractLinearOperator,
    FunctionLinearOperator,
    IdentityLinearOperator,
    conj,
    has_unit_diagonal,
    is_diagonal,
    is_lower_triangular,
    is_negative_semidefinite,
    is_positive_semidefinite,
    is_symmetric,
    is_tridiagonal,
    is_upper_triangular,
)
from ._solution import RESULTS, Solution
from ._structure import (
    batch_shape_of,
    inexact_asarray,
    ravel,
    tree_leaves,
    tree_map,
)
from ._tags import (
    diagonal_tag,
    lower_triangular_tag,
    negative_semidefinite_tag,
    positive_semidefinite_tag,
    symmetric_tag,
    unit_diagonal_tag,
    upper_triangular_tag,
)


class _Sentinel:
    def __repr__(self):
        return "sentinel"


sentinel = _Sentinel()


class AbstractLinearSolver(abc.ABC):
    """Abstract base class for linear solvers (lineax/_solve.py:343-480).

    ``init(operator, options) -> state`` does the work that depends only on
    the operator (e.g. a factorisation); ``compute(state, vector, options)``
    returns ``(solution, result, stats)``; ``transpose``/``conj`` map a state
    to the state of the transposed/conjugated operator without re-factoring;
    ``assume_full_rank`` selects the cheaper differentiation rule.
    """

    @abc.abstractmethod
    def init(self, operator: AbstractLinearOperator, options: dict[str, Any]):
        ...

    @abc.abstractmethod
    def compute(self, state, vector, options: dict[str, Any]):
        ...

    @abc.abstractmethod
    def transpose(self, state, options: dict[str, Any]):
        ...

    @abc.abstractmethod
    def conj(self, state, options: dict[str, Any]):
        ...

    @abc.abstractmethod
    def assume_full_rank(self) -> bool:
        ...

    def __eq__(self, other):
        return type(self) is type(other) and self.__dict__ == other.__dict__

    def __hash__(self):
        return hash((type(self), tuple(sorted((k, repr(v)) for k, v in self.__dict__.items()))))

    def __repr__(self):
        fields = ", ".join(f"{k}={v!r}" for k, v in self.__dict__.items())
        return f"{type(self).__name__}({fields})"


# ---------------------------------------------------------------------------
# The primitive: compute + RESULTS post-processing
# ---------------------------------------------------------------------------


def _result_tensor(result, batch: tuple, device) -> torch.Tensor:
    r = torch.as_tensor(result, dtype=torch.int32, device=device)
    return r.expand(batch).contiguous().clone() if r.shape != tuple(batch) else r.contiguous().clone()


def _linear_solve_impl(operator, state, vector, options, solver, throw: bool):
    solution, result, stats = solver.compute(state, vector, options)
    vleaves = tree_leaves(vector)
    sleaves = tree_leaves(solution)
    device = vleaves[0].device if vleaves else None
    batch = batch_shape_of(solution, operator.in_structure())
    vbatch = batch_shape_of(vector, operator.out_structure())
    result = _result_tensor(result, batch, device)
    if sleaves and device is not None and device.type == "cuda":
        flat_s = ravel(solution, operator.in_structure(), batch)
        flat_v = ravel(vector, operator.out_structure(), vbatch)
        flat_v = flat_v.expand(*batch, flat_v.shape[-1])
        if flat_s.shape[-1] != flat_v.shape[-1]:
            # Non-square: scan the two vectors separately.
            K.solve_results(flat_s, torch.zeros_like(flat_s), result)
            bad_in = ~torch.isfinite(flat_v).all(dim=-1)
            result.copy_(torch.where((result == RESULTS.singular) & bad_in, RESULTS.nonfinite_input, result))
        else:
            K.solve_results(flat_s, flat_v.to(flat_s.dtype), result)
    if throw:
        bad = result != RESULTS.successful
        if bool(bad.any()):
            raise RuntimeError(RESULTS[result])
    return solution, result, stats


# ---------------------------------------------------------------------------
# Autodiff attachment
# ---------------------------------------------------------------------------


def _needs_grad(tensors) -> bool:
    import torch.autograd.forward_ad as fwAD

    for t in tensors:
        if not isinstance(t, torch.Tensor):
            continue
        if torch.is_grad_enabled() and t.requires_grad:
            return True
        if fwAD.unpack_dual(t).tangent is not None:
            return True
    return False


def _detach(tree):
    return tree_map(lambda t: t.detach() if isinstance(t, torch.Tensor) else t, tree)


def _detach_operator(op: AbstractLinearOperator) -> AbstractLinearOperator:
    leaves = op._leaves()
    if not leaves:
        return op
    return op._with_leaves([t.detach() for t in leaves])


class _SolveHandle:
    """Everything a derivative solve needs: the operator, state and solver."""

    def __init__(self, operator, state, options, solver, out_shape, dtype):
        self.operator = operator
        self.state = state
        self.options = options
        self.solver = solver
        self.out_shape = out_shape
        self.dtype = dtype

    def solve(self, vec):
        x, _, _ = _linear_solve_impl(self.operator, self.state, vec, self.options, self.solver, True)
        return x

    def solve_transpose(self, vec):
        state_t, options_t = self.solver.transpose(self.state, self.options)
        x, _, _ = _linear_solve_impl(self.operator.transpose(), state_t, vec, options_t, self.solver, True)
        return x


class _ZeroSolve(torch.autograd.Function):
    """Value zero; derivative ``handle.solve`` (forward) / ``handle.solve_transpose`` (reverse)."""

    @staticmethod
    def forward(r, handle):
        return torch.zeros(handle.out_shape, dtype=handle.dtype, device=r.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.handle = inputs[1]
        ctx.r_shape = inputs[0].shape

    @staticmethod
    def backward(ctx, g):
        ct = ctx.handle.solve_transpose(g.contiguous())
        return ct.reshape(ctx.r_shape) if ct.shape != ctx.r_shape else ct, None

    @staticmethod
    def jvp(ctx, dr, _):
        return ctx.handle.solve(dr.contiguous())


def _attach_derivatives(operator, state, vector, options, solver, solution):
    """x = x* + Z(b - A x*) (+ pseudoinverse terms for non-square operators)."""
    if not isinstance(vector, torch.Tensor) or not isinstance(solution, torch.Tensor):
        raise NotImplementedError("differentiating through linear_solve needs single-tensor vectors")
    x_star = solution.detach()
    op_d = _detach_operator(operator)
    solve_a = _SolveHandle(op_d, state, options, solver, x_star.shape, x_star.dtype)
    rows, cols = operator.out_size(), operator.in_size()
    # A^+ (b' - A' x)
    out = x_star + _ZeroSolve.apply(vector - operator.mv(x_star), solve_a)
    independent_rows = solver.assume_full_rank() and rows <= cols
    independent_cols = solver.assume_full_rank() and cols <= rows
    if independent_rows and independent_cols:
        return out
    # lineax/_solve.py:216-253 (real dtypes, so conj is the identity).
    op_t = conj(operator).transpose()
    op_t_d = _detach_operator(op_t)
    state_c, options_c = solver.conj(state, options)
    state_ct, options_ct = solver.transpose(state_c, options_c)
    batch = tuple(x_star.shape[:-1])
    if not independent_rows:
        # + A^+ A^+T A'^T (b - A x): the tangent of q = A^T r* mapped by
        # A^+T (solve with the transposed state) and then by A^+.
        resid = (vector - operator.mv(x_star)).detach()
        q = op_t.mv(resid)
        solve_at = _SolveHandle(op_t_d, state_ct, options_ct, solver, (*batch, rows), x_star.dtype)
        out = out + _ZeroSolve.apply(_ZeroSolve.apply(q, solve_at), solve_a)
    if not independent_cols:
        # + y' - A^+ A y'  with  y = A'^T A^+T x.
        h, _, _ = _linear_solve_impl(op_t_d, state_ct, x_star, options_ct, solver, True)
        y = op_t.mv(h.detach())
        y_t = y - y.detach()  # value 0, tangent A'^T h
        out = out + y_t + _ZeroSolve.apply(-op_d.mv(y_t), solve_a)
    return out


# ---------------------------------------------------------------------------
# AutoLinearSolver
# ---------------------------------------------------------------------------


class AutoLinearSolver(AbstractLinearSolver):
    """Picks a solver from the operator's structure (lineax/_solve.py:518-645).

    ``well_posed=True``: diagonal -> Diagonal, tridiagonal -> Tridiagonal,
    triangular -> Triangular, PSD/NSD -> Cholesky, else LU.
    ``well_posed=None``: non-square -> QR, then as above.
    ``well_posed=False``: diagonal -> Diagonal, else SVD.
    """

    def __init__(self, well_posed: bool | None):
        self.well_posed = well_posed

    def _select(self, operator) -> str:
        square = operator.in_size() == operator.out_size()
        if self.well_posed is True:
            if not square:
                raise ValueError(
                    "Cannot use `AutoLinearSolver(well_posed=True)` with a non-square operator. If you "
                    "are trying solve a least-squares problem then you should pass "
                    "`solver=AutoLinearSolver(well_posed=False)`. By default `linear_solve` assumes "
                    "that the operator is square and nonsingular."
                )
            return self._square_choice(operator, well_posed_diagonal=True)
        if self.well_posed is False:
            return "diagonal" if is_diagonal(operator) else "svd"
        if self.well_posed is None:
            if not square:
                return "qr"
            return self._square_choice(operator, well_posed_diagonal=False)
        raise ValueError(f"Invalid value `well_posed={self.well_posed}`.")

    @staticmethod
    def _square_choice(operator, well_posed_diagonal: bool) -> str:
        if is_diagonal(operator):
            return "well_posed_diagonal" if well_posed_diagonal else "diagonal"
        if is_tridiagonal(operator):
            return "tridiagonal"
        if is_lower_triangular(operator) or is_upper_triangular(operator):
            return "triangular"
        if is_positive_semidefinite(operator) or is_negative_semidefinite(operator):
            return "cholesky"
        return "lu"

    def select_solver(self, operator) -> AbstractLinearSolver:
        return _lookup(self._select(operator))

    def init(self, operator, options):
        token = self._select(operator)
        return token, _lookup(token).init(operator, options)

    def compute(self, state, vector, options):
        token, inner = state
        solution, result, _ = _lookup(token).compute(inner, vector, options)
        return solution, result, {}

    def transpose(self, state, options):
        token, inner = state
        st, opt = _lookup(token).transpose(inner, options)
        return (token, st), opt

    def conj(self, state, options):
        token, inner = state
        st, opt = _lookup(token).conj(inner, options)
        return (token, st), opt

    def assume_full_rank(self):
        return self.well_posed is not False


def _lookup(token: str) -> AbstractLinearSolver:
    from . import _solver

    table = {
        "qr": _solver.QR,
        "cholesky": _solver.Cholesky,
        "lu": _solver.LU,
        "triangular": _solver.Triangular,
        "diagonal": lambda: _solver.Diagonal(),
        "well_posed_diagonal": lambda: _solver.Diagonal(well_posed=True),
        "tridiagonal": _solver.Tridiagonal,
        "svd": _solver.SVD,
    }
    return table[token]()


# ---------------------------------------------------------------------------
# linear_solve
# ---------------------------------------------------------------------------


def linear_solve(
    operator: AbstractLinearOperator,
    vector,
    solver: AbstractLinearSolver | None = None,
    *,
    options: dict[str, Any] | None = None,
    state: Any = sentinel,
    throw: bool = True,
) -> Solution:
    """Solves ``operator @ x = vector`` (lineax/_solve.py:656-806).

    Returns the pseudoinverse solution: the usual solution for square
    nonsingular operators, least squares for overdetermined and minimum norm
    for underdetermined ones (depending on the solver).

    - ``operator``: an ``AbstractLinearOperator`` (wrap matrices in
      ``MatrixLinearOperator``).  Leading batch dimensions on the operator
      and/or the vector solve a batch of independent systems in one launch.
    - ``solver``: defaults to ``AutoLinearSolver(well_posed=True)``.
    - ``options``: solver-specific runtime options (e.g. ``preconditioner``,
      ``y0`` for CG/GMRES).
    - ``state``: the result of ``solver.init(operator, options)``, to reuse a
      factorisation across right-hand sides.
    - ``throw``: raise on failure (``True``) or report it in
      ``Solution.result`` (``False``).  Raising synchronises with the device.
    """
    if isinstance(operator, torch.Tensor):
        raise ValueError(
            "`linear_solve(operator=...)` should be an `AbstractLinearOperator`, not a raw tensor. "
            "If you are trying to pass a matrix then this should be passed as "
            "`MatrixLinearOperator(matrix)`."
        )
    if solver is None:
        solver = AutoLinearSolver(well_posed=True)
    if options is None:
        options = {}
    vector = tree_map(inexact_asarray, vector)
    batch_shape_of(vector, operator.out_structure())  # raises on a structure mismatch
    if isinstance(operator, IdentityLinearOperator):
        vleaves = tree_leaves(vector)
        batch = batch_shape_of(vector, operator.out_structure())
        dev = vleaves[0].device if vleaves else None
        return Solution(
            value=vector,
            result=torch.zeros(batch, dtype=torch.int32, device=dev),
            state=state,
            stats={},
        )
    if state is sentinel:
        with torch.no_grad():
            state = solver.init(_detach_operator(operator), options)
    state = _detach(state)

    ad = _needs_grad(tree_leaves(vector) + operator._leaves())
    if not ad:
        with torch.no_grad():
            solution, result, stats = _linear_solve_impl(operator, state, vector, options, solver, throw)
    else:
        with torch.no_grad():
            solution, result, stats = _linear_solve_impl(
                _detach_operator(operator), state, _detach(vector), options, solver, throw
            )
        solution = _attach_derivatives(operator, state, vector, options, solver, solution)
    return Solution(value=solution, result=result, state=state, stats=stats)


def invert(operator, solver: AbstractLinearSolver | None = None, *, options=None, throw: bool = True):
    """A ``FunctionLinearOperator`` whose ``mv`` solves with ``operator``.

    The factorisation is computed once (lineax/_solve.py:809-871).
    """
    if solver is None:
        solver = AutoLinearSolver(well_posed=True)
    if options is None:
        options = {}
    state = solver.init(operator, options)

    def solve_fn(vector):
        return linear_solve(operator, vector, solver, state=state, options=options, throw=throw).value

    tags = {
        tag
        for check, tag in (
            (is_symmetric, symmetric_tag),
            (is_diagonal, diagonal_tag),
            (is_lower_triangular, lower_triangular_tag),
            (is_upper_triangular, upper_triangular_tag),
            (is_positive_semidefinite, positive_semidefinite_tag),
            (is_negative_semidefinite, negative_semidefinite_tag),
        )
        if check(operator)
    }
    if has_unit_diagonal(operator) and (
        is_diagonal(operator) or is_lower_triangular(operator) or is_upper_triangular(operator)
    ):
        tags.add(unit_diagonal_tag)
    return FunctionLinearOperator(
        solve_fn,
        operator.out_structure(),
        frozenset(tags),
        output_structure=operator.in_structure(),
        device=operator.device,
    )
