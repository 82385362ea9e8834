"""CG solver (lineax/_solver/cg.py) driving the device Krylov kernels."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._misc import resolve_rcond
from .._norm import max_norm, rms_norm, two_norm
from .._operator import (
    IdentityLinearOperator,
    NegLinearOperator,
    TaggedLinearOperator,
    conj,
    mv_as,
    mv_into,
    is_negative_semidefinite,
    is_positive_semidefinite,
    linearise,
)
from .._solve import AbstractLinearSolver
from .._structure import batch_shape_of, ravel, structure_equal, unravel
from .misc import preconditioner_and_y0

NORM_KINDS = {max_norm: 0, two_norm: 1, rms_norm: 2}
# Iterations between host checks of the per-row `active` flags: grows from
# 1 to this, so short solves stop promptly and long ones sync rarely.
MAX_CHECK_BLOCK = 32
# Iterations per captured graph (rounded up to a multiple of stabilise_every).
GRAPH_BLOCK = 30


def stencil_of(operator):
    """(nx, ny, coeffs) when ``operator`` is a five-point stencil operator,
    possibly tagged or negated; else None."""
    from ..stencil import StencilOperator

    sign = 1.0
    while True:
        if isinstance(operator, TaggedLinearOperator):
            operator = operator.operator
        elif isinstance(operator, NegLinearOperator):
            operator = operator.operator
            sign = -sign
        else:
            break
    if isinstance(operator, StencilOperator):
        return operator.nx, operator.ny, tuple(sign * c for c in operator.coeffs)
    return None


_FUSED_DTYPES = {(torch.float64, torch.float32), (torch.float64, torch.float64), (torch.float32, torch.float32)}


def norm_kind(norm) -> int:
    try:
        return NORM_KINDS[norm]
    except KeyError:
        raise NotImplementedError(
            "lineax_amd Krylov solvers support norm=max_norm, two_norm or rms_norm"
        ) from None


class FlatOperator:
    """An operator's matvec on flat [rows, n] tensors, in a chosen output dtype."""

    def __init__(self, operator, struct, batch):
        self.op = operator
        self.struct = struct
        self.identity = isinstance(operator, IdentityLinearOperator)

    def __call__(self, v: torch.Tensor, out_dtype: torch.dtype) -> torch.Tensor:
        if self.identity:
            return v if v.dtype == out_dtype else v.to(out_dtype)
        return mv_as(self.op, v, self.struct, out_dtype)

    def into(self, v: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """The product written into ``out`` (flat, contiguous, storage dtype)."""
        if self.identity:
            return out.copy_(v)
        return mv_into(self.op, v, self.struct, out)


class CG(AbstractLinearSolver):
    """Conjugate gradients for positive or negative definite operators.

    Semantics of lineax/_solver/cg.py:48-250: stabilised residual every
    ``stabilise_every`` steps, the dual stopping test in both the residual and
    the iterate, the ``rcond`` breakdown guard, max_steps = 10 * size by
    default and the RESULTS mapping.  ``working_dtype`` (not in lineax)
    stores r, p and A p in a narrower dtype than the vector's: with a float64
    right-hand side and ``working_dtype=torch.float32`` the iterate, the
    stabilising residuals and all inner products stay float64.  ``graph``
    (not in lineax) captures blocks of iterations in a HIP graph and replays
    them, checking convergence on the host between replays; the operator
    must then be capturable (device work only, no host synchronisation).
    ``fused`` (not in lineax): for a five-point stencil operator with no
    preconditioner, run each iteration as the two fused sweeps of
    lxa_cg_stencil_direction / _update instead of six separate kernels.
    """

    def __init__(self, rtol, atol, norm=max_norm, stabilise_every: int | None = 10, max_steps: int | None = None,
                 working_dtype: torch.dtype | None = None, graph: bool = False, fused: bool = True):
        if isinstance(rtol, (int, float)) and rtol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if isinstance(atol, (int, float)) and atol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if atol == 0 and rtol == 0 and max_steps is None:
            raise ValueError("Must specify `rtol`, `atol`, or `max_steps` (or some combination of all three).")
        self.rtol = rtol
        self.atol = atol
        self.norm = norm
        self.stabilise_every = stabilise_every
        self.max_steps = max_steps
        self.working_dtype = working_dtype
        self.graph = graph
        self.fused = fused

    def init(self, operator, options):
        is_nsd = is_negative_semidefinite(operator)
        if not structure_equal(operator.in_structure(), operator.out_structure()):
            raise ValueError("`CG()` may only be used for linear solves with square matrices.")
        if not (is_positive_semidefinite(operator) or is_nsd):
            raise ValueError("`CG()` may only be used for positive or negative definite linear operators")
        if is_nsd:
            operator = -operator
        return linearise(operator), is_nsd

    def compute(self, state, vector, options):
        operator, is_nsd = state
        preconditioner, y0 = preconditioner_and_y0(operator, vector, options)
        if not is_positive_semidefinite(preconditioner):
            raise ValueError("The preconditioner must be positive definite.")
        struct = operator.in_structure()
        batch = batch_shape_of(vector, struct)
        b = ravel(vector, struct, batch)
        size = b.shape[-1]
        rows = b.numel() // size if size else 0
        b = b.reshape(rows, size).contiguous()
        ydt = b.dtype
        vdt = self.working_dtype or ydt
        dev = b.device
        max_steps = 10 * size if self.max_steps is None else int(self.max_steps)
        prm = K.KrylovParams(
            rtol=float(self.rtol), atol=float(self.atol), rcond=resolve_rcond(None, size, size, ydt), eps=0.0,
            max_steps=max_steps, max_steps_given=int(self.max_steps is not None), norm_kind=norm_kind(self.norm),
            restart=0, stagnation_iters=0,
        )
        A = FlatOperator(operator, struct, batch)
        M = FlatOperator(preconditioner, struct, batch)
        precond = not M.identity
        yd, vd = K.dtype_code(ydt), K.dtype_code(vdt)
        st = K.RowState(rows, dev)
        part = K.partials_buffer(rows, size, 0, dev)
        stream = K.current_stream(b)

        y = ravel(y0, struct, batch).reshape(rows, size).to(ydt).contiguous().clone()
        r = torch.empty((rows, size), dtype=vdt, device=dev)
        st.active().fill_(1)
        t = A(y, ydt)
        K.call("lxa_cg_residual", stream, yd, vd, rows, size, b, t, r, 0, prm, st, part)
        z = M(r, vdt) if precond else r
        p = z.clone()
        K.call("lxa_cg_begin", stream, vd, rows, size, z, r, prm, st, part)

        se = self.stabilise_every
        stencil = stencil_of(operator) if (self.fused and not precond and dev.type == "cuda") else None
        if stencil is not None and (stencil[0] * stencil[1] != size or (ydt, vdt) not in _FUSED_DTYPES):
            stencil = None

        if stencil is None:
            def iteration(i):
                stream = K.current_stream(b)  # the capture stream while a graph is recorded
                q = A(p, vdt)
                K.call("lxa_cg_alpha", stream, vd, rows, size, q, p, prm, st, part)
                stable = se == 1 or (se is not None and (i + 1) % se == 0)
                fuse = (not precond) and (not stable)
                K.call("lxa_cg_update", stream, yd, vd, rows, size, p, q, y, r, b, int(stable), int(fuse), prm, st,
                       part)
                if stable:
                    t = A(y, ydt)
                    K.call("lxa_cg_residual", stream, yd, vd, rows, size, b, t, r, int(not precond), prm, st, part)
                if precond:
                    z = M(r, vdt)
                    K.call("lxa_cg_precond_gamma", stream, yd, vd, rows, size, b, z, r, prm, st, part)
                else:
                    z = r
                K.call("lxa_cg_finish", stream, vd, rows, size, z, p, prm, st, part)
        else:
            nx, ny, coeffs = stencil
            ws = K.StencilCGWorkspace(rows, nx, ny, coeffs, dev)
            # b is read in float32 when that is exact (the usual case of float32
            # data promoted to a float64 solve): same values, half the traffic
            b_read, bd = b, yd
            if ydt == torch.float64 and vdt == torch.float32:
                b32 = b.to(torch.float32)
                if torch.equal(b32.to(torch.float64), b):
                    b_read, bd = b32, K.dtype_code(torch.float32)
            pbuf = (p, torch.empty_like(p))
            # stabilising steps write the iterate to the other buffer
            ybuf = [y, torch.empty_like(y)]
            ycur = [0]

            def iteration(i):
                stream = K.current_stream(b)
                src, dst = pbuf[i % 2], pbuf[(i + 1) % 2]
                K.call("lxa_cg_stencil_direction", stream, vd, rows, nx, ny, ws.coeffs, r, src, dst, int(i == 0), prm,
                       st, ws.partials)
                stable = se == 1 or (se is not None and (i + 1) % se == 0)
                y_in = ybuf[ycur[0]]
                if stable:
                    ycur[0] ^= 1
                y_out = ybuf[ycur[0]]
                K.call("lxa_cg_stencil_update", stream, yd, vd, bd, rows, nx, ny, ws.coeffs, dst, y_in, y_out, r,
                       b_read, int(stable), prm, st, ws.partials)

        it = 0
        block = 1
        # Graph mode: a block of `period` iterations (a multiple of the
        # stabilisation interval, so every replay has the same schedule) is
        # captured once and replayed between host checks of the active flags.
        period = 0
        if self.graph and dev.type == "cuda":
            unit = se or 1
            if stencil is not None:
                # the fused sweeps alternate two direction buffers each step and
                # two iterate buffers each stabilising step: a replayed block
                # holds an even number of both
                unit = 2 * (se or 1)
            period = unit * max(1, -(-GRAPH_BLOCK // unit))
        graph = None
        while True:
            if graph is None and period and it % period == 0 and it + period <= max_steps and it >= period:
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    for j in range(period):
                        iteration(it + j)
            if graph is not None and it + period <= max_steps:
                graph.replay()
                it += period
            else:
                for _ in range(block):
                    iteration(it)
                    it += 1
                    if it >= max_steps or (period and it % period == 0):
                        break
            if it >= max_steps or not bool(st.active().any()):
                break
            block = min(block * 2, MAX_CHECK_BLOCK)
        stream = K.current_stream(b)
        K.call("lxa_cg_result", stream, rows, prm, st)

        if stencil is not None:
            y = ybuf[ycur[0]]
        if is_nsd:
            y = -y
        solution = unravel(y.reshape(*batch, size), struct, batch)
        result = st.result().clone().reshape(batch)
        stats = {"num_steps": st.step().clone().reshape(batch), "max_steps": self.max_steps}
        return solution, result, stats

    def transpose(self, state, options):
        transpose_options = {}
        if "preconditioner" in options:
            transpose_options["preconditioner"] = options["preconditioner"].transpose()
        psd_op, is_nsd = state
        return (psd_op.transpose(), is_nsd), transpose_options

    def conj(self, state, options):
        conj_options = {}
        if "preconditioner" in options:
            conj_options["preconditioner"] = conj(options["preconditioner"])
        psd_op, is_nsd = state
        return (conj(psd_op), is_nsd), conj_options

    def assume_full_rank(self):
        return True
