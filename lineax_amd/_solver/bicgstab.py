"""BiCGStab solver (lineax/_solver/bicgstab.py) driving the device Krylov kernels."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._norm import max_norm
from .._operator import conj, linearise
from .._solve import AbstractLinearSolver
from .._structure import batch_shape_of, ravel, structure_equal, unravel
from .cg import GRAPH_BLOCK, MAX_CHECK_BLOCK, FlatOperator, norm_kind
from .misc import preconditioner_and_y0

# Breakdown threshold of lineax without jax_enable_x64 (bicgstab.py:112-113);
# float64 solves test for exact zeros (bicgstab.py:110-111).
F32_BREAKDOWN = 1e-16


class BiCGStab(AbstractLinearSolver):
    """Biconjugate gradient stabilised method with right preconditioning.

    Semantics of lineax/_solver/bicgstab.py:35-222: the iteration of
    jax.scipy.sparse.linalg.bicgstab (body 143-167), the dual stopping test
    in the residual and the update (cond_fun 121-140), the breakdown test on
    omega, alpha and rho, max_steps = 10 * size by default and the RESULTS
    mapping (185-197).  The breakdown test follows the solve's dtype: exact
    zeros for float64 (lineax under jax_enable_x64), values below 1e-16 for
    float32.  ``working_dtype`` and ``graph`` as for CG (not in lineax).
    """

    def __init__(self, rtol, atol, norm=max_norm, max_steps: int | None = None,
                 working_dtype: torch.dtype | None = None, graph: bool = False):
        if isinstance(rtol, (int, float)) and rtol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if isinstance(atol, (int, float)) and atol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if atol == 0 and rtol == 0 and max_steps is None:
            raise ValueError("Must specify `rtol`, `atol`, or `max_steps` (or some combination of all three).")
        self.rtol = rtol
        self.atol = atol
        self.norm = norm
        self.max_steps = max_steps
        self.working_dtype = working_dtype
        self.graph = graph

    def init(self, operator, options):
        if not structure_equal(operator.in_structure(), operator.out_structure()):
            raise ValueError("`BiCGstab(..., normal=False)` may only be used for linear solves with square matrices.")
        return linearise(operator)

    def compute(self, state, vector, options):
        operator = state
        preconditioner, y0 = preconditioner_and_y0(operator, vector, options)
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
            rtol=float(self.rtol), atol=float(self.atol), rcond=0.0,
            eps=0.0 if ydt == torch.float64 else F32_BREAKDOWN, max_steps=max_steps,
            max_steps_given=int(self.max_steps is not None), norm_kind=norm_kind(self.norm), restart=0,
            stagnation_iters=0,
        )
        A = FlatOperator(operator, struct, batch)
        M = FlatOperator(preconditioner, struct, batch)
        precond = not M.identity
        yd, vd = K.dtype_code(ydt), K.dtype_code(vdt)
        st = K.RowState(rows, dev)
        part = K.partials_buffer(rows, size, 0, dev)

        y = ravel(y0, struct, batch).reshape(rows, size).to(ydt).contiguous().clone()
        r = torch.empty((rows, size), dtype=vdt, device=dev)
        r0 = torch.empty_like(r)
        p = torch.empty_like(r)
        v = torch.zeros_like(r)  # carried across iterations: a fixed buffer (graph replays read it)
        s = torch.empty_like(r)
        t0 = A(y, ydt)
        K.call("lxa_bicgstab_begin", K.current_stream(b), yd, vd, rows, size, b, t0, r, r0, prm, st, part)

        def iteration(i):
            stream = K.current_stream(b)  # the capture stream while a graph is recorded
            K.call("lxa_bicgstab_direction", stream, vd, rows, size, r, p, v, int(i == 0), st)
            x = M(p, vdt) if precond else p
            A.into(x, v)
            K.call("lxa_bicgstab_alpha", stream, vd, rows, size, r0, v, prm, st, part)
            K.call("lxa_bicgstab_s", stream, vd, rows, size, r, v, s, st)
            z = M(s, vdt) if precond else s
            t = A(z, vdt)
            K.call("lxa_bicgstab_omega", stream, vd, rows, size, s, t, prm, st, part)
            K.call("lxa_bicgstab_update", stream, yd, vd, rows, size, x, z, s, t, r0, y, r, b, prm, st, part)

        it = 0
        block = 1
        period = GRAPH_BLOCK if (self.graph and dev.type == "cuda") else 0
        graph = None
        while it < max_steps and bool(st.active().any()):
            if graph is None and period and it >= period and it + period <= max_steps:
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
            block = min(block * 2, MAX_CHECK_BLOCK)
        K.call("lxa_bicgstab_result", K.current_stream(b), rows, prm, st)

        solution = unravel(y.reshape(*batch, size), struct, batch)
        result = st.result().clone().reshape(batch)
        stats = {"num_steps": st.step().clone().reshape(batch), "max_steps": self.max_steps}
        return solution, result, stats

    def transpose(self, state, options):
        transpose_options = {}
        if "preconditioner" in options:
            transpose_options["preconditioner"] = options["preconditioner"].transpose()
        return state.transpose(), transpose_options

    def conj(self, state, options):
        conj_options = {}
        if "preconditioner" in options:
            conj_options["preconditioner"] = conj(options["preconditioner"])
        return conj(state), conj_options

    def assume_full_rank(self):
        return True
