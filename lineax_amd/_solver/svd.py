"""SVD solver (lineax/_solver/svd.py) on the one-sided Jacobi SVD kernel."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._misc import resolve_rcond
from .._operator import matvec
from .._solution import RESULTS
from .._solve import AbstractLinearSolver
from .misc import pack_structures, ravel_vector, transpose_packed_structures, unravel_solution


class SVD(AbstractLinearSolver):
    """Pseudoinverse solve through a thin SVD, for any operator (nonsquare or
    singular), as scipy.linalg.lstsq.

    ``init`` factors the operator's matrix (lineax/_solver/svd.py:49-53);
    ``compute`` keeps the singular values above rcond * s_max (rcond defaults
    to 2 eps max(m, n); lineax/_solver/svd.py:55-78) and returns
    V diag(1/s) U^T b with the rank in the stats; ``transpose`` swaps the
    factors (svd.py:80-93).  State: ``((u, s, vt), packed_structures,
    converged)``; a Jacobi iteration that did not converge within its sweep
    budget (lineax's LAPACK SVD reports nothing) gives
    ``RESULTS.max_steps_reached`` rather than passing inexact factors off as
    successful.
    """

    def __init__(self, rcond: float | None = None):
        self.rcond = rcond

    def init(self, operator, options):
        u, s, vt, sweeps = K.svd(operator.as_matrix())
        converged = sweeps <= K.SVD_MAX_SWEEPS
        return (u, s, vt), pack_structures(operator), converged

    def compute(self, state, vector, options):
        (u, s, vt), packed, converged = state
        flat, batch = ravel_vector(vector, packed)
        flat = flat.to(u.dtype)
        m, n = u.shape[-2], vt.shape[-1]
        rcond = resolve_rcond(self.rcond, n, m, s.dtype)
        cutoff = rcond * s[..., :1] if s.shape[-1] > 0 else torch.zeros_like(s[..., :1])
        # strictly greater, so an all-zero matrix has rank 0 (svd.py:70-71)
        mask = s > cutoff
        rank = mask.sum(dim=-1)
        s_inv = torch.where(mask, 1.0 / torch.where(mask, s, torch.ones_like(s)), torch.zeros_like(s))
        utb = matvec(u.mT, flat)
        solution = matvec(vt.mT, s_inv * utb)
        result = torch.where(converged, RESULTS.successful, RESULTS.max_steps_reached).to(torch.int32)
        return unravel_solution(solution, packed, tuple(solution.shape[:-1])), result, {"rank": rank}

    def transpose(self, state, options):
        (u, s, vt), packed, converged = state
        return ((vt.mT, s, u.mT), transpose_packed_structures(packed), converged), {}

    def conj(self, state, options):
        (u, s, vt), packed, converged = state
        return ((u.conj(), s, vt.conj()), packed, converged), {}

    def assume_full_rank(self):
        return False
