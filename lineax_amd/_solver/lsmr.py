"""LSMR solver (lineax/_solver/lsmr.py) driving the device LSMR kernels."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._norm import two_norm
from .._operator import conj, linearise
from .._solve import AbstractLinearSolver
from .._structure import batch_shape_of, ravel, tree_leaves, unravel
from .cg import GRAPH_BLOCK, MAX_CHECK_BLOCK, FlatOperator


class LSMR(AbstractLinearSolver):
    """Least-squares solver of Fong and Saunders for A x = b, any shape of A.

    Semantics of lineax/_solver/lsmr.py:55-441 with damp = 0: the
    Golub-Kahan bidiagonalisation, the rotations Qhat, Q and Qbar, the
    estimates of |r|, |A^T r|, |A| and cond(A) (224-306), the stopping rules
    in their order (309-325), max_steps = 10 * min(m, n) by default and the
    RESULTS mapping (347-362).  Each batch row runs its own recurrence: the
    vectors u, v, h, hbar and x stay in device memory and the per-row
    scalars in a device state array, updated by the lxa_lsmr_* kernels; the
    host only applies A and A^T and polls the rows' active flags.

    Only the two-norm is provided (lineax's default and the norm the
    recurrence estimates).  ``working_dtype`` (not in lineax) stores u, v, h
    and hbar in a narrower dtype than x; the scalars are float64 throughout.
    ``graph`` (not in lineax) captures blocks of iterations in a HIP graph
    and replays them, checking the rows' active flags on the host between
    replays; the operator must then be capturable (device work only).
    """

    def __init__(self, rtol, atol, norm=two_norm, max_steps: int | None = None, conlim: float = 1e8,
                 working_dtype: torch.dtype | None = None, graph: bool = False):
        for value in (rtol, atol, conlim):
            if isinstance(value, (int, float)) and value < 0:
                raise ValueError("Tolerances must be non-negative.")
        if atol == 0 and rtol == 0 and max_steps is None:
            raise ValueError("Must specify `atol`, `rtol`, or `max_steps` (or some combination of all three).")
        if norm is not two_norm:
            raise NotImplementedError("lineax_amd LSMR supports norm=two_norm only")
        self.rtol = rtol
        self.atol = atol
        self.norm = norm
        self.max_steps = max_steps
        self.conlim = conlim
        self.working_dtype = working_dtype
        self.graph = graph

    def init(self, operator, options):
        return linearise(operator)

    def compute(self, state, vector, options):
        operator = state
        in_struct, out_struct = operator.in_structure(), operator.out_structure()
        batch = batch_shape_of(vector, out_struct)
        b = ravel(vector, out_struct, batch)
        m = b.shape[-1]
        n = sum(leaf.size for leaf in tree_leaves(in_struct))
        rows = b.numel() // m if m else 0
        b = b.reshape(rows, m).contiguous()
        ydt = b.dtype
        vdt = self.working_dtype or ydt
        dev = b.device
        if dev.type != "cuda":
            raise ValueError("LSMR runs on the GPU: the vector must be a device tensor")
        max_steps = min(m, n) * 10 if self.max_steps is None else int(self.max_steps)
        prm = K.LsmrParams(rtol=float(self.rtol), atol=float(self.atol), conlim=float(self.conlim),
                           max_steps=max_steps, max_steps_given=int(self.max_steps is not None), pad=0)
        A = FlatOperator(operator, in_struct, batch)
        At = FlatOperator(operator.transpose(), out_struct, batch)
        yd, vd = K.dtype_code(ydt), K.dtype_code(vdt)
        st = K.LsmrState(rows, dev)
        part = K.partials_buffer(rows, max(m, n), 0, dev)

        if "y0" in options:
            x = ravel(options["y0"], in_struct, batch).reshape(rows, n).to(ydt).contiguous().clone()
        else:
            x = torch.zeros((rows, n), dtype=ydt, device=dev)
        u = torch.empty((rows, m), dtype=vdt, device=dev)
        v = torch.empty((rows, n), dtype=vdt, device=dev)
        h = torch.empty_like(v)
        hbar = torch.empty_like(v)
        stream = K.current_stream(b)
        # u = b - A x0 (unscaled), beta = |u|, |b|
        K.call("lxa_lsmr_begin", stream, yd, vd, rows, m, b, A(x, ydt).contiguous(), u, prm, st, part)
        # v = A^T u / beta, alpha = |v|, h = v / alpha, hbar = 0, istop for beta = 0 or alpha = 0
        K.call("lxa_lsmr_init", stream, vd, rows, n, At(u, vdt).contiguous(), v, h, hbar, prm, st, part)

        def iteration():
            stream = K.current_stream(b)  # the capture stream while a graph is recorded
            K.call("lxa_lsmr_step_u", stream, vd, rows, m, A(v, vdt).contiguous(), u, prm, st, part)
            K.call("lxa_lsmr_step_v", stream, vd, rows, n, At(u, vdt).contiguous(), v, prm, st, part)
            K.call("lxa_lsmr_step_x", stream, yd, vd, rows, n, v, h, hbar, x, prm, st, part)

        # Rows that have stopped skip every kernel (their active flag is
        # read on the device), so whole blocks may run past a row's stop.
        it = 0
        block = 1
        period = GRAPH_BLOCK if (self.graph and dev.type == "cuda") else 0
        graph = None
        while it < max_steps and bool(st.int_field("active").any()):
            if graph is None and period and it >= period and it + period <= max_steps:
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    for _ in range(period):
                        iteration()
            if graph is not None and it + period <= max_steps:
                graph.replay()
                it += period
            else:
                for _ in range(block):
                    iteration()
                    it += 1
                    if it >= max_steps or (period and it % period == 0):
                        break
            block = min(block * 2, MAX_CHECK_BLOCK)
        K.call("lxa_lsmr_result", K.current_stream(b), rows, prm, st)

        solution = unravel(x.reshape(*batch, n), in_struct, batch)
        result = st.int_field("result").clone().reshape(batch)
        stats = {
            "num_steps": st.int_field("itn").clone().reshape(batch),
            "istop": st.int_field("istop").clone().reshape(batch),
            "norm_r": st.scalar("normr").clone().reshape(batch),
            "norm_Ar": st.scalar("normAr").clone().reshape(batch),
            "norm_A": st.scalar("normA2").sqrt().reshape(batch),
            "cond_A": st.scalar("condA").clone().reshape(batch),
            "norm_x": torch.linalg.vector_norm(x.double(), dim=-1).reshape(batch),
            "max_steps": self.max_steps,
        }
        return solution, result, stats

    def transpose(self, state, options):
        del options
        transpose_options = {}
        return state.transpose(), transpose_options

    def conj(self, state, options):
        del options
        conj_options = {}
        return conj(state), conj_options

    def assume_full_rank(self):
        return False
