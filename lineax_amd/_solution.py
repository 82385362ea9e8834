"""RESULTS codes and the Solution container (lineax/_solution.py:52-88).

Codes are integers in lineax's declaration order; per-system results are
int32 device tensors (0-d for an unbatched solve, the batch shape when the
operator or vector carries leading batch dimensions).
"""

from __future__ import annotations

import dataclasses
from typing import Any

import torch

_SINGULAR_MSG = (
    "A linear solver returned non-finite (NaN or inf) output. This usually means that an\n"
    "operator was not well-posed, and that its solver does not support this.\n\n"
    "If you are trying solve a linear least-squares problem then you should pass\n"
    "`solver=AutoLinearSolver(well_posed=False)`. By default `linear_solve`\n"
    "assumes that the operator is square and nonsingular.\n\n"
    "If you *were* expecting this solver to work with this operator, then it may be because:\n\n"
    "(a) the operator is singular, and your code has a bug; or\n\n"
    "(b) the operator was nearly singular (i.e. it had a high condition number), and the\n"
    "    solver suffered from numerical instability issues; or\n\n"
    "(c) the operator is declared to exhibit a certain property (e.g. positive definiteness)\n"
    "    that is does not actually satisfy."
)

_NONFINITE_MSG = (
    "A linear solver received non-finite (NaN or inf) input and cannot determine a\n"
    "solution.\n\n"
    "This means that you have a bug upstream and should check the inputs to\n"
    "`linear_solve` for non-finite values."
)


class _ResultsMeta(type):
    def __getitem__(cls, code) -> str:
        """Message for a result code (int, 0-d tensor, or a batch: first failure)."""
        if isinstance(code, torch.Tensor):
            flat = code.reshape(-1)
            if flat.numel() == 0:
                return cls._messages[0]
            bad = flat[flat != 0]
            code = int(bad[0]) if bad.numel() else 0
        return cls._messages[int(code)]

    def __iter__(cls):
        return iter(cls._names)

    def __len__(cls):
        return len(cls._names)


class RESULTS(metaclass=_ResultsMeta):
    """Integer result codes of a solve, in lineax's order."""

    successful = 0
    max_steps_reached = 1
    singular = 2
    breakdown = 3
    stagnation = 4
    conlim = 5
    nonfinite_input = 6

    _names = (
        "successful",
        "max_steps_reached",
        "singular",
        "breakdown",
        "stagnation",
        "conlim",
        "nonfinite_input",
    )
    _messages = (
        "",
        "The maximum number of solver steps was reached. Try increasing `max_steps`.",
        _SINGULAR_MSG,
        "A form of iterative breakdown has occured in a linear solve. Try using a different "
        "solver for this problem or increase `restart` if using GMRES.",
        "A stagnation in an iterative linear solve has occurred. Try increasing "
        "`stagnation_iters` or `restart`.",
        "Condition number of A seems to be larger than `conlim`.",
        _NONFINITE_MSG,
    )

    @staticmethod
    def where(pred, if_true, if_false):
        """Elementwise select between result codes (device tensors or ints)."""
        pred = torch.as_tensor(pred)
        dev = pred.device
        t = torch.as_tensor(if_true, dtype=torch.int32, device=dev)
        f = torch.as_tensor(if_false, dtype=torch.int32, device=dev)
        return torch.where(pred, t, f).to(torch.int32)

    @staticmethod
    def name(code: int) -> str:
        return RESULTS._names[int(code)]


@dataclasses.dataclass
class Solution:
    """The solution to a linear solve (lineax/_solution.py:71-88).

    - ``value``: the solution (a tensor or pytree of tensors).
    - ``result``: int32 result code(s); ``RESULTS[result]`` gives the message.
    - ``stats``: solver statistics, e.g. ``num_steps`` for iterative solvers.
    - ``state``: the solver state used, reusable via ``linear_solve(..., state=)``.
    """

    value: Any
    result: torch.Tensor
    stats: dict
    state: Any
