"""A request loop that solves a stream of dense least-squares (QR) and SPD
(Cholesky) systems on one GPU with continuous batching.

Requests arrive over time (``Request.arrival``, seconds after the start).
The chosen form (``SolveServer(form="chosen")``) runs in ticks:

1. admission: arrived requests are admitted first-come first-served while
   their working memory fits in the HBM arena (``hbm_budget`` bytes; the
   operator, its factors, scratch and right-hand side are allocated there);
2. ``panels_per_tick`` launches of ``lxa_qr_factor_step`` advance every
   admitted QR system by that many 16-column panels, each at its own column,
   so systems admitted in different ticks share the launches (continuous
   batching);
3. systems whose factorisation completed, and requests that reuse a resident
   factorisation with a new right-hand side, are solved together (one
   ``lxa_qr_apply`` + ``lxa_trsv`` launch pair);
4. newly admitted Cholesky systems are factored and solved in one uneven
   launch pair (``lxa_chol_blocked_factor`` / ``_solve``);
5. the float32 solutions of the tick are refined in FP64 (``refine_steps``
   steps, batched over the tick's requests): Cholesky requests by classical
   refinement (y = b - A x in FP64, correction with the float32 factor),
   least-squares requests by the corrected semi-normal equations
   (x += R^-1 R^-T A^T (b - A x), products in FP64, triangular solves with
   the float32 R), which contract the error by about kappa(A) u_f32 per step;
6. completions are found by polling events; requests completed in a tick get
   that completion time.  Finished factorisations stay resident (up to
   ``resident_budget`` bytes, least recently used evicted) for reuse
   requests.

The plain form (``form="plain"``) is static batching with one launch per
size bucket: when idle it takes every request that has arrived as one batch,
groups it by kind and size bucket, runs the float32 kernels once per bucket
(uneven QR factor + apply + trsv, uneven blocked Cholesky), without refinement
or residency, and takes the next batch when this one is done.  ``form="fifo"`` runs the same kernels one request at a time in
arrival order.  Their answers are float32.

Semantically each request is ``lx.linear_solve(lx.MatrixLinearOperator(A),
b, lx.QR())`` (lineax/_solver/qr.py:55-95; least squares for m >= n) or
``lx.linear_solve(lx.MatrixLinearOperator(A, positive_semidefinite_tag), b,
lx.Cholesky())`` (lineax/_solver/cholesky.py:43-78) on float32 data; the
chosen form returns the refined solution in float64.
"""

from __future__ import annotations

import bisect
import dataclasses
import time

import numpy as np
import torch

from . import _kernels as K

QR_ELEM = 4  # float32 bytes
SMALL_CHOL = 256  # Cholesky orders served by the uneven tile-class launch
PANELS_PER_TICK = 16  # QR panel steps per tick of the factor stream (swept: 4, 8, 16)
PLAIN_QR_BUCKET = 256  # rows per size bucket of the plain form's least-squares launches
PLAIN_CHOL_BUCKET = 64  # orders per size bucket of the plain form's Cholesky launches
SOLVE_STREAMS = 2  # with the factor stream and the default stream: the 4 hardware queues of a process


@dataclasses.dataclass
class Request:
    rid: int
    kind: str  # "qr" or "chol"
    arrival: float
    m: int
    n: int
    op_id: int  # operator identity; reuse requests share an earlier request's op_id
    reuse: bool
    a: torch.Tensor | None  # the operator (device, float32), None for reuse requests
    b: torch.Tensor  # right-hand side (device, float32, length m)
    # filled in by the server
    x: torch.Tensor | None = None
    admitted: float | None = None
    done: float | None = None


class Arena:
    """First-fit allocator over one device buffer (offsets in float32 elements)."""

    ALIGN = 64  # elements (256 bytes)

    def __init__(self, elems: int, device):
        self.size = int(elems)
        self.buf = torch.empty(self.size, dtype=torch.float32, device=device)
        self.free = [(0, self.size)]  # sorted (offset, length)
        self.used = 0
        self.high_water = 0

    def alloc(self, elems: int) -> int | None:
        need = max(self.ALIGN, -(-int(elems) // self.ALIGN) * self.ALIGN)
        for i, (off, ln) in enumerate(self.free):
            if ln >= need:
                if ln == need:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + need, ln - need)
                self.used += need
                self.high_water = max(self.high_water, self.used)
                return off
        return None

    def release(self, off: int, elems: int) -> None:
        need = max(self.ALIGN, -(-int(elems) // self.ALIGN) * self.ALIGN)
        self.used -= need
        i = bisect.bisect_left(self.free, (off, 0))
        self.free.insert(i, (off, need))
        # coalesce with neighbours
        if i + 1 < len(self.free) and self.free[i][0] + self.free[i][1] == self.free[i + 1][0]:
            self.free[i] = (self.free[i][0], self.free[i][1] + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)

    def view(self, off: int, elems: int) -> torch.Tensor:
        return self.buf[off: off + elems]


def qr_footprint(m: int, n: int, nb: int = 16) -> dict:
    """Arena elements of one QR system: matrix, tau, T blocks, scratch, rhs."""
    k = min(m, n)
    return {"a": m * n, "tau": k, "t": -(-k // nb) * nb * nb, "w": -(-m // 256) * nb * n, "b": m}


def chol_footprint(n: int) -> dict:
    return {"a": n * n, "b": n}


class _Resident:
    """A finished factorisation kept in the arena for reuse requests."""

    def __init__(self, kind, m, n, blocks):
        self.kind, self.m, self.n, self.blocks = kind, m, n, blocks
        self.lda = n
        self.last_use = 0
        self.busy = 0  # reuse solves in flight


class SolveServer:
    def __init__(self, device, *, form: str = "chosen", hbm_budget: int = 8 << 30,
                 resident_budget: int = 2 << 30, refine_steps: int = 2, panels_per_tick: int = PANELS_PER_TICK):
        if form not in ("chosen", "plain", "fifo"):
            raise ValueError("form must be 'chosen', 'plain' or 'fifo'")
        if refine_steps < 0:
            raise ValueError("refine_steps must be >= 0")
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.form = form
        self.refine_steps = int(refine_steps) if form == "chosen" else 0
        self.arena = Arena(hbm_budget // 4, self.device)
        self.resident_budget = resident_budget // 4
        self.resident: dict[int, _Resident] = {}
        self.resident_elems = 0
        self.nb = K.qr_block(torch.float32)
        if panels_per_tick < 1:
            raise ValueError("panels_per_tick must be >= 1")
        self.panels_per_tick = int(panels_per_tick)
        # the tile-class launch keeps its work-queue counters in this scratch:
        # one per solve stream, since launches on different streams overlap
        # (a shared one lets one launch take the other's queue entries)
        self._uneven_ws = [torch.zeros(256, dtype=torch.uint8, device=self.device) for _ in range(SOLVE_STREAMS)]
        self.ticks = 0
        self.tick_log: list[dict] = []

    # ------------------------------------------------------------------ memory
    def _alloc_blocks(self, sizes: dict) -> dict | None:
        got = {}
        for key, elems in sizes.items():
            off = self.arena.alloc(elems)
            if off is None:
                for k2, (o2, e2) in got.items():
                    self.arena.release(o2, e2)
                return None
            got[key] = (off, elems)
        return got

    def _free_blocks(self, blocks: dict) -> None:
        for off, elems in blocks.values():
            self.arena.release(off, elems)

    def _evict_one(self) -> bool:
        idle = [k for k, r in self.resident.items() if r.busy == 0]
        if not idle:
            return False
        victim = min(idle, key=lambda k: self.resident[k].last_use)
        r = self.resident.pop(victim)
        self.resident_elems -= sum(e for _, e in r.blocks.values())
        self._free_blocks(r.blocks)
        return True

    def _evict_until(self, need_elems: int) -> None:
        while self.resident_elems + need_elems > self.resident_budget:
            idle = [k for k, r in self.resident.items() if r.busy == 0]
            if not idle:
                break
            victim = min(idle, key=lambda k: self.resident[k].last_use)
            r = self.resident.pop(victim)
            self.resident_elems -= sum(e for _, e in r.blocks.values())
            self._free_blocks(r.blocks)

    # ------------------------------------------------------------------ kernels
    def _descs(self, items):
        """items: (m, n, lda, blocks)."""
        d = np.zeros(len(items), dtype=K.QR_DESC)
        for i, (m, n, lda, blocks) in enumerate(items):
            d[i]["a_offset"] = blocks["a"][0]
            if "tau" in blocks:
                d[i]["tau_offset"] = blocks["tau"][0]
                d[i]["t_offset"] = blocks["t"][0]
                if "w" in blocks:
                    d[i]["w_offset"] = blocks["w"][0]
            d[i]["b_offset"] = blocks["b"][0]
            d[i]["m"], d[i]["n"], d[i]["lda"], d[i]["ldb"] = m, n, lda, m
        return K.descs_to_device(d, self.device)

    def _qr_solve_batch(self, items):
        """items: (m, n, lda, blocks) with the rhs in blocks['b']; x = first n entries."""
        if not items:
            return
        buf = self.arena.buf
        dd = self._descs(items)
        max_m = max(it[0] for it in items)
        max_k = max(min(it[0], it[1]) for it in items)
        K.qr_apply_raw(dd, len(items), max_m, max_k, buf, buf, 1, buf, trans=True)
        K.trsv_raw(dd, len(items), max_k, buf, 1, buf, lower=False, trans=False, unit_diag=False)

    def _trsv_batch(self, items):
        """R x = c for factored systems whose rhs blocks hold c = (Q^T b)[:n]."""
        if not items:
            return
        buf = self.arena.buf
        dd = self._descs(items)
        max_k = max(min(it[0], it[1]) for it in items)
        K.trsv_raw(dd, len(items), max_k, buf, 1, buf, lower=False, trans=False, unit_diag=False)

    def _footprint(self, r: Request) -> dict:
        """Arena elements of a new request: in the chosen form QR factors the
        augmented [A | b] and keeps an n-vector for the solution."""
        if self.form == "chosen" and r.kind == "qr":
            fp = qr_footprint(r.m, r.n + 1, self.nb)
            fp["b"] = r.n
            return fp
        return qr_footprint(r.m, r.n, self.nb) if r.kind == "qr" else chol_footprint(r.n)

    def _chol_batch(self, items, factor: bool):
        if not items:
            return
        buf = self.arena.buf
        dd = self._descs(items)
        if factor:
            info = torch.empty(len(items), dtype=torch.int32, device=self.device)
            K.chol_blocked_factor_raw(dd, len(items), buf, info)
        K.chol_blocked_solve_raw(dd, len(items), buf, buf)

    def _solutions(self, batch):
        """The answers of a tick's solves, batch = [(request, item, rhs blocks)]
        with the float32 solution in the first n entries of each item's b
        block.  Without refinement: float32 copies.  With refinement: FP64
        solutions after ``refine_steps`` batched steps, all on the current
        stream with no host synchronisation:
          y = b - A x (FP64, every request);
          Cholesky:       x += (U^T U)^-1 y;
          least squares:  x += R^-1 R^-T (A^T y)  (corrected semi-normal
                          equations with the float32 R; A^T y in FP64).
        Each step contracts the error (in the norm |R e|) by about
        2 kappa(A) u_f32 plus the triangular solves' rounding, so Gaussian
        least-squares problems with kappa up to ~1e5 reach the FP64 level in
        two to three steps without applying Q."""
        buf = self.arena.buf
        if self.refine_steps == 0:
            return [buf[it[3]["b"][0]: it[3]["b"][0] + r.n].clone() for r, it, _ in batch]
        dev = self.device
        order = ([i for i, e in enumerate(batch) if e[0].kind == "qr"]
                 + [i for i, e in enumerate(batch) if e[0].kind != "qr"])
        ents = [batch[i] for i in order]
        nq = sum(1 for e in ents if e[0].kind == "qr")
        cnt = len(ents)
        m = np.array([e[0].m for e in ents], dtype=np.int64)
        n = np.array([e[0].n for e in ents], dtype=np.int64)
        offx = np.concatenate(([0], np.cumsum(n)[:-1]))
        offm = np.concatenate(([0], np.cumsum(m)[:-1]))
        nt, mt = int(n.sum()), int(m.sum())
        nq_el, mq_el = int(n[:nq].sum()), int(m[:nq].sum())
        bo = np.array([e[1][3]["b"][0] for e in ents], dtype=np.int64)
        # arena positions of the solutions: the first n entries of each b block
        idx = K.host_to_device(torch.from_numpy(np.arange(nt) + np.repeat(bo - offx, n)), dev)
        x = buf[idx].double()
        y = torch.empty(mt, dtype=torch.float64, device=dev)
        g = torch.empty(nq_el, dtype=torch.float64, device=dev)
        # FP64 products: y = b - A x over all requests; g = A^T y over the QR ones
        gd = np.zeros(cnt + nq, dtype=K.GEMV_DESC)
        ops = [e[0].a if e[0].a is not None else self._operator_data[e[0].op_id] for e in ents]
        aptr = np.array([a.data_ptr() for a in ops], dtype=np.int64)
        gd["a"][:cnt] = aptr
        gd["x"][:cnt] = x.data_ptr() + 8 * offx
        gd["b"][:cnt] = [e[0].b.data_ptr() for e in ents]
        gd["y"][:cnt] = y.data_ptr() + 8 * offm
        gd["m"][:cnt], gd["n"][:cnt], gd["lda"][:cnt] = m, n, n
        gd["a"][cnt:] = aptr[:nq]
        gd["x"][cnt:] = y.data_ptr() + 8 * offm[:nq]
        gd["y"][cnt:] = g.data_ptr() + 8 * offx[:nq]
        gd["m"][cnt:], gd["n"][cnt:], gd["lda"][cnt:] = m[:nq], n[:nq], n[:nq]
        # factor descriptors (QR first, then Cholesky), right-hand side in the b block
        qd = np.zeros(cnt, dtype=K.QR_DESC)
        for i, (r, (mi, ni, lda, blk), _) in enumerate(ents):
            qd[i]["a_offset"] = blk["a"][0]
            if "tau" in blk:
                qd[i]["tau_offset"], qd[i]["t_offset"] = blk["tau"][0], blk["t"][0]
            qd[i]["b_offset"] = blk["b"][0]
            qd[i]["m"], qd[i]["n"], qd[i]["lda"], qd[i]["ldb"] = mi, ni, lda, mi
        raw = K.host_to_device(torch.from_numpy(np.concatenate((gd.view(np.uint8), qd.view(np.uint8)))), dev)
        g0, g1 = raw[: cnt * 48], raw[cnt * 48: (cnt + nq) * 48]
        d_q = raw[(cnt + nq) * 48: (cnt + nq) * 48 + nq * 56]
        d_c = raw[(cnt + nq) * 48 + nq * 56:]
        max_m, max_n = int(m.max()), int(n.max())
        max_mq = int(m[:nq].max()) if nq else 0
        max_kq = int(n[:nq].max()) if nq else 0
        for _ in range(self.refine_steps):
            K.gemv_uneven_raw(g0, cnt, False, max_m, max_n, buf)
            if nq:
                K.gemv_uneven_raw(g1, nq, True, max_mq, max_kq, buf)
            # right-hand sides of the corrections, in the solution slots
            buf[idx] = torch.cat((g, y[mq_el:])).float() if nq else y.float()
            if nq:
                K.trsv_raw(d_q, nq, max_kq, buf, 1, buf, lower=False, trans=True, unit_diag=False)
                K.trsv_raw(d_q, nq, max_kq, buf, 1, buf, lower=False, trans=False, unit_diag=False)
            if cnt > nq:
                K.chol_blocked_solve_raw(d_c, cnt - nq, buf, buf)
            x += buf[idx].double()
        xs = [None] * cnt
        for j, i in enumerate(order):
            xs[i] = x[int(offx[j]): int(offx[j]) + int(n[j])]
        return xs

    # ------------------------------------------------------------------ run
    def run(self, requests: list[Request], operators: dict, *, time_limit: float = 120.0) -> dict:
        """Serves the stream in real time; returns latency and throughput figures.

        ``operators`` maps op_id to the operator's device matrix (reuse
        requests carry no matrix of their own)."""
        self._operator_data = operators
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            a = r.a if r.a is not None else operators.get(r.op_id)
            if a is None or a.shape != (r.m, r.n) or a.dtype != torch.float32 or not a.is_contiguous():
                raise ValueError(f"request {r.rid}: the operator must be a contiguous float32 {r.m}x{r.n} tensor")
            if r.b.shape != (r.m,) or r.b.dtype != torch.float32 or not r.b.is_contiguous():
                raise ValueError(f"request {r.rid}: b must be a contiguous float32 vector of length {r.m}")
            if a.device != self.device or r.b.device != self.device:
                raise ValueError(f"request {r.rid}: operator and b must live on {self.device}")
            if r.kind == "qr" and r.m < r.n:
                raise ValueError("the request loop serves least-squares QR requests (m >= n)")
            fp = self._footprint(r)
            if sum(fp.values()) + len(fp) * Arena.ALIGN > self.arena.size:
                raise MemoryError(f"request {r.rid} ({r.kind} {r.m}x{r.n}) exceeds the HBM arena")
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        if self.form == "fifo":
            self._run_plain(reqs, t0, time_limit)
        elif self.form == "plain":
            self._run_static(reqs, t0, time_limit)
        else:
            self._run_chosen(reqs, t0, time_limit)
        makespan = time.perf_counter() - t0
        lat = np.array([r.done - r.arrival for r in reqs if r.done is not None])
        done = int(len(lat))
        return {
            "requests": len(reqs),
            "completed": done,
            "makespan_s": makespan,
            "solves_per_s": done / makespan if makespan > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if done else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if done else None,
            "ticks": self.ticks,
            "arena_high_water_bytes": self.arena.high_water * 4,
        }

    def _now(self, t0):
        return time.perf_counter() - t0

    def _run_plain(self, reqs, t0, time_limit):
        for r in reqs:
            while self._now(t0) < r.arrival:
                pass
            if self._now(t0) > time_limit:
                break
            r.admitted = self._now(t0)
            if r.kind == "qr":
                fp = qr_footprint(r.m, r.n, self.nb)
                blocks = self._alloc_blocks(fp)
                if blocks is None:
                    raise MemoryError("request does not fit in the arena")
                a = r.a if r.a is not None else self._operator_data[r.op_id]
                self.arena.view(*blocks["a"]).copy_(a.reshape(-1))
                self.arena.view(*blocks["b"]).copy_(r.b)
                dd = self._descs([(r.m, r.n, r.n, blocks)])
                buf = self.arena.buf
                K.qr_factor_raw(dd, 1, r.m, r.n, min(r.m, r.n), buf, buf, buf, buf)
                self._qr_solve_batch([(r.m, r.n, r.n, blocks)])
                r.x = self.arena.view(blocks["b"][0], r.n).clone()
            else:
                fp = chol_footprint(r.n)
                blocks = self._alloc_blocks(fp)
                if blocks is None:
                    raise MemoryError("request does not fit in the arena")
                a = r.a if r.a is not None else self._operator_data[r.op_id]
                self.arena.view(*blocks["a"]).copy_(a.reshape(-1))
                self.arena.view(*blocks["b"]).copy_(r.b)
                self._chol_batch([(r.n, r.n, r.n, blocks)], factor=True)
                r.x = self.arena.view(*blocks["b"]).clone()
            torch.cuda.synchronize(self.device)
            r.done = self._now(t0)
            self._free_blocks(blocks)
            self.ticks += 1

    def _run_static(self, reqs, t0, time_limit):
        """The plain form: static batching with one launch per size bucket.
        When idle, every request that has arrived (as many as fit in the
        arena) becomes one batch.  Its requests are grouped by kind and size
        bucket (least squares by m in PLAIN_QR_BUCKET-row buckets, Cholesky by
        n in PLAIN_CHOL_BUCKET buckets), and each bucket runs the FP32 kernels
        once: an uneven lxa_qr_factor launch sequence and one apply + trsv
        pair, or one lxa_chol_blocked factor + solve pair.  No refinement and
        no residency (a reuse request factors its operator again); the next
        batch starts when this one is done."""
        i = 0
        buf = self.arena.buf
        while i < len(reqs):
            now = self._now(t0)
            if now > time_limit:
                break
            if reqs[i].arrival > now:
                continue
            qr_items, ch_items, taken = [], [], []
            while i < len(reqs) and reqs[i].arrival <= now:
                r = reqs[i]
                fp = qr_footprint(r.m, r.n, self.nb) if r.kind == "qr" else chol_footprint(r.n)
                blocks = self._alloc_blocks(fp)
                if blocks is None:
                    if not taken:
                        raise MemoryError("request does not fit in the arena")
                    break
                a = r.a if r.a is not None else self._operator_data[r.op_id]
                self.arena.view(*blocks["a"]).copy_(a.reshape(-1))
                self.arena.view(*blocks["b"]).copy_(r.b)
                r.admitted = now
                (qr_items if r.kind == "qr" else ch_items).append((r.m, r.n, r.n, blocks))
                taken.append((r, blocks))
                i += 1
            qr_buckets: dict[int, list] = {}
            for it in qr_items:
                qr_buckets.setdefault((it[0] - 1) // PLAIN_QR_BUCKET, []).append(it)
            for _, items in sorted(qr_buckets.items()):
                dd = self._descs(items)
                max_m = max(it[0] for it in items)
                max_n = max(it[1] for it in items)
                K.qr_factor_raw(dd, len(items), max_m, max_n, min(max_m, max_n), buf, buf, buf, buf)
                self._qr_solve_batch(items)
            ch_buckets: dict[int, list] = {}
            for it in ch_items:
                ch_buckets.setdefault((it[0] - 1) // PLAIN_CHOL_BUCKET, []).append(it)
            for _, items in sorted(ch_buckets.items()):
                self._chol_batch(items, factor=True)
            xs = [self.arena.view(blocks["b"][0], r.n).clone() for r, blocks in taken]
            torch.cuda.synchronize(self.device)
            t_done = self._now(t0)
            for (r, blocks), x in zip(taken, xs):
                r.x, r.done = x, t_done
                self._free_blocks(blocks)
            self.ticks += 1

    def _run_chosen(self, reqs, t0, time_limit):
        """Continuous batching on a factor stream and SOLVE_STREAMS solve streams.

        Factor stream: one batched panel step per tick for every QR system being
        factored.  A new QR request carries its right-hand side as an extra
        column of A (leading dimension n + 1), so the panel updates leave Q^T b
        in that column and only the triangular solve remains.  Solve streams
        (one per tick, in turn): triangular solves of finished QR systems
        (after an event on the factor stream), Cholesky factor + solve of new
        requests, reuse solves, and the refinement of all of them.  A tick's
        solve chain is latency-bound (one workgroup per system), so chains of
        consecutive ticks run side by side on different compute units.
        Completions are found by polling events; arena blocks are returned only
        after the work that uses them has finished.  At most `depth` ticks are
        in flight on the factor stream."""
        dev = self.device
        fstream = torch.cuda.Stream(device=dev)
        sstreams = [torch.cuda.Stream(device=dev) for _ in range(SOLVE_STREAMS)]
        depth = 2
        pending = list(reqs)
        waiting: list[Request] = []
        factoring: list[list] = []  # [request, blocks, next column]
        owner_pending: dict[int, Request] = {}
        inflight: list[tuple] = []  # (event, [(request, x, item, rhs_blocks)])
        tick_events: list = []
        i_next = 0
        buf = self.arena.buf

        def finish_batches(block: bool) -> None:
            # chains on different solve streams may finish out of order
            if block and inflight:
                inflight[0][0].synchronize()
            done = [e for e in inflight if e[0].query()]
            if not done:
                return
            t_done = self._now(t0)
            for e in done:
                inflight.remove(e)
                for r, x, item, rhs_blocks in e[1]:
                    r.x, r.done = x, t_done
                    self._retire(r, item, rhs_blocks, owner_pending)

        while True:
            now = self._now(t0)
            if now > time_limit:
                break
            finish_batches(False)
            while i_next < len(pending) and pending[i_next].arrival <= now:
                waiting.append(pending[i_next])
                i_next += 1
            if not waiting and not factoring:
                if i_next >= len(pending):
                    break
                if inflight:
                    finish_batches(False)
                continue
            qr_done, chol_new, solve_reuse, chol_small = [], [], [], []
            still, blocked = [], False
            si = self.ticks % len(sstreams)
            sstream = sstreams[si]
            with torch.cuda.stream(fstream):
                for r in waiting:
                    if blocked:
                        still.append(r)
                        continue
                    if r.kind == "chol" and r.n <= SMALL_CHOL and self.refine_steps > 0:
                        # factored, solved and refined in one persistent launch
                        # straight from the operator's storage (no arena copy);
                        # unrefined serving keeps the float32 potrf batches
                        r.admitted = now
                        chol_small.append(r)
                        continue
                    if r.reuse and r.op_id in self.resident:
                        res = self.resident[r.op_id]
                        blocks = self._alloc_blocks({"b": r.m})
                        if blocks is None:
                            blocked = True
                            still.append(r)
                            continue
                        with torch.cuda.stream(sstream):
                            self.arena.view(*blocks["b"]).copy_(r.b)
                        res.last_use = self.ticks
                        res.busy += 1
                        merged = dict(res.blocks)
                        merged.update(blocks)
                        solve_reuse.append((r, (r.m, r.n, res.lda, merged), blocks))
                        continue
                    if r.reuse and r.op_id in owner_pending:
                        still.append(r)
                        continue
                    a = r.a if r.a is not None else self._operator_data[r.op_id]
                    blocks = self._alloc_blocks(self._footprint(r))
                    if blocks is None:
                        blocked = True
                        still.append(r)
                        continue
                    r.admitted = now
                    owner_pending[r.op_id] = r
                    if r.kind == "qr":
                        aug = self.arena.view(*blocks["a"]).view(r.m, r.n + 1)
                        aug[:, : r.n].copy_(a)
                        aug[:, r.n].copy_(r.b)
                        factoring.append([r, blocks, 0])
                    else:
                        with torch.cuda.stream(sstream):
                            self.arena.view(*blocks["a"]).copy_(a.reshape(-1))
                            self.arena.view(*blocks["b"]).copy_(r.b)
                        chol_new.append((r, (r.n, r.n, r.n, blocks), None))
                waiting = still
                if blocked and not (factoring or qr_done or chol_new or solve_reuse or chol_small or inflight):
                    if not self._evict_one():
                        raise MemoryError("admission blocked with nothing in flight")
                    continue
                # one panel step for every QR system being factored (factor stream)
                if factoring:
                    items = [(r.m, r.n + 1, r.n + 1, blk) for r, blk, _ in factoring]
                    dd = self._descs(items)
                    # pinned, non-blocking (a pageable copy would wait for the stream)
                    kcol = K.host_to_device(torch.tensor([k for _, _, k in factoring], dtype=torch.int32), dev)
                    max_rows = max(r.m - k for r, _, k in factoring)
                    max_tc = max(r.n - k for r, _, k in factoring)
                    # several panel steps per tick: the host work of a tick
                    # (descriptors, events, polling) is paid once for them;
                    # systems past their last column are skipped by the kernels
                    steps = min(self.panels_per_tick, -(-max(r.n - k for r, _, k in factoring) // self.nb))
                    for j in range(steps):
                        if j:
                            kcol += self.nb
                        K.qr_factor_step_raw(dd, len(factoring), kcol, max_rows, max_tc, buf, buf, buf, buf)
                    nxt = []
                    for f in factoring:
                        f[2] += self.nb * steps
                        (qr_done if f[2] >= f[0].n else nxt).append(f)
                    factoring = nxt
                ev_f = torch.cuda.Event()
                ev_f.record(fstream)
            # solves (solve stream)
            batch = []
            with torch.cuda.stream(sstream):
                if qr_done:
                    sstream.wait_event(ev_f)
                    items = []
                    for r, blk, _ in qr_done:
                        # Q^T b sits in column n of the augmented matrix
                        col = self.arena.buf[blk["a"][0] + r.n: blk["a"][0] + r.n + r.n * (r.n + 1): r.n + 1]
                        self.arena.view(*blk["b"]).copy_(col)
                        items.append((r.m, r.n, r.n + 1, blk))
                    self._trsv_batch(items)
                    for (r, blk, _), item in zip(qr_done, items):
                        batch.append((r, item, None))
                qr_reuse = [(r, it, rb) for r, it, rb in solve_reuse if r.kind == "qr"]
                ch_reuse = [(r, it, rb) for r, it, rb in solve_reuse if r.kind != "qr"]
                if qr_reuse:
                    self._qr_solve_batch([it for _, it, _ in qr_reuse])
                self._chol_batch([it for _, it, _ in chol_new], factor=True)
                self._chol_batch([it for _, it, _ in ch_reuse], factor=False)
                batch += qr_reuse + chol_new + ch_reuse
                if batch:
                    xs = self._solutions(batch)
                    batch = [(r, x, item, rb) for (r, item, rb), x in zip(batch, xs)]
                if chol_small:
                    batch += [(r, x, None, None) for r, x in zip(chol_small, self._chol_small_solve(chol_small, si))]
                if batch:
                    ev = torch.cuda.Event()
                    ev.record(sstream)
                    inflight.append((ev, batch))
            self.ticks += 1
            tick_events.append(ev_f)
            if len(tick_events) > depth:
                tick_events.pop(0).synchronize()
        finish_batches(True)
        while inflight:
            finish_batches(True)

    def _chol_small_solve(self, reqs, stream_slot: int = 0):
        """Cholesky requests of order <= SMALL_CHOL: one lxa_chol_uneven_solve
        (persistent tile-class launches on the matrix cores, FP64 refinement to
        eta <= 1e-12) reading each operator where it lives.  Returns float64
        solutions, on the current stream.  stream_slot picks the work-queue
        scratch of the solve stream the call runs on: calls on different
        streams must not share one."""
        dev = self.device
        ops = [r.a if r.a is not None else self._operator_data[r.op_id] for r in reqs]
        ptr = np.array([a.data_ptr() for a in ops], dtype=np.int64)
        base = int(ptr.min())
        base_t = next(a for a in ops if a.data_ptr() == base)
        n = np.array([r.n for r in reqs], dtype=np.int64)
        b_off = np.concatenate(([0], np.cumsum(n)[:-1]))
        order = np.lexsort((np.arange(len(reqs)), -n))
        # one host-to-device copy: a offsets, b offsets, orders, work order
        meta = np.concatenate(((ptr - base) // 4, b_off, n, order)).astype(np.int64)
        meta_d = K.host_to_device(torch.from_numpy(meta), dev)
        cnt = len(reqs)
        a_off_d, b_off_d = meta_d[:cnt], meta_d[cnt: 2 * cnt]
        n_d, order_d = meta_d[2 * cnt: 3 * cnt].to(torch.int32), meta_d[3 * cnt:].to(torch.int32)
        b = torch.cat([r.b for r in reqs])
        x = torch.empty(b.numel(), dtype=torch.float64, device=dev)
        info = torch.empty(cnt, dtype=torch.int32, device=dev)
        a_all = torch.as_strided(base_t, (1,), (1,))  # base pointer; offsets reach every operator
        K.chol_uneven_solve(n_d, a_off_d, a_all, b_off_d, b, x, order_d, max_n=int(n.max()), persistent=True,
                            max_refine=10, tol=1e-12, info=info, workspace=self._uneven_ws[stream_slot],
                            sorted_order=True)
        return [x[int(o): int(o) + int(k)] for o, k in zip(b_off, n)]

    def _retire(self, r, item, rhs_blocks, owner_pending):
        """Returns a finished request's blocks to the arena or keeps its
        factorisation resident for reuse requests."""
        if item is None:  # a small Cholesky request: nothing in the arena
            if owner_pending.get(r.op_id) is r:
                owner_pending.pop(r.op_id, None)
            return
        m, n, lda, blk = item
        if rhs_blocks is not None:
            self._free_blocks(rhs_blocks)  # a reuse solve: the resident factor stays
            self.resident[r.op_id].busy -= 1
            return
        owner_pending.pop(r.op_id, None)
        # keep what a reuse solve needs: the factor (and tau, T blocks for QR)
        keep = {k: v for k, v in blk.items() if k not in ("b", "w")}
        for k in ("b", "w"):
            if k in blk:
                self.arena.release(*blk[k])
        elems = sum(e for _, e in keep.values())
        if elems <= self.resident_budget:
            self._evict_until(elems)
            if r.op_id in self.resident:  # an older copy
                old = self.resident.pop(r.op_id)
                self.resident_elems -= sum(e for _, e in old.blocks.values())
                self._free_blocks(old.blocks)
            res = _Resident(r.kind, m, n, keep)
            res.lda = lda
            res.last_use = self.ticks
            self.resident[r.op_id] = res
            self.resident_elems += elems
        else:
            self._free_blocks(keep)


def make_request_stream(rate: float, count: int, *, seed: int = 5, device="cuda:0", qr_m=(64, 4096),
                        chol_n=(16, 1024), reuse_fraction: float = 0.1,
                        qr_fraction: float = 0.5) -> tuple[list[Request], dict]:
    """Synthetic Poisson stream: exponential gaps at `rate`/s; half QR
    (m ~ U{qr_m}, n ~ U{ceil(m/4)..m}, N(0,1) entries), half SPD Cholesky
    (n ~ U{chol_n}, A = M M^T / n + 0.1 I); a `reuse_fraction` of requests
    reuse an earlier operator with a new right-hand side.  Data is generated
    on the device (seeded per request) before serving starts.  Returns the
    requests and the operators by op_id."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs: list[Request] = []
    ops: dict[int, tuple[str, int, int]] = {}
    data: dict[int, torch.Tensor] = {}
    gen = torch.Generator(device=device)
    for i in range(count):
        reuse = bool(ops) and rng.random() < reuse_fraction
        if reuse:
            op_id = int(rng.choice(list(ops)))
            kind, m, n = ops[op_id]
            a = None
        else:
            op_id = i
            if rng.random() < qr_fraction:
                kind = "qr"
                m = int(rng.integers(qr_m[0], qr_m[1] + 1))
                n = int(rng.integers(-(-m // 4), m + 1))
            else:
                kind = "chol"
                n = int(rng.integers(chol_n[0], chol_n[1] + 1))
                m = n
            ops[op_id] = (kind, m, n)
            gen.manual_seed(seed * 1_000_003 + i)
            if kind == "qr":
                a = torch.randn((m, n), generator=gen, device=device, dtype=torch.float32)
            else:
                mm = torch.randn((n, n), generator=gen, device=device, dtype=torch.float64)
                a = (mm @ mm.T / n + 0.1 * torch.eye(n, device=device, dtype=torch.float64)).to(torch.float32)
            data[op_id] = a
        gen.manual_seed(seed * 1_000_003 + 500_000 + i)
        b = torch.randn((m,), generator=gen, device=device, dtype=torch.float32)
        reqs.append(Request(rid=i, kind=kind, arrival=float(arrivals[i]), m=m, n=n, op_id=op_id, reuse=reuse,
                            a=a, b=b))
    return reqs, data


def request_flops(r: Request) -> float:
    """Algorithmic flops of one request (a reuse request: the solve only)."""
    if r.kind == "qr":
        m, n = r.m, r.n
        solve = 4.0 * m * n
        return solve if r.reuse else 2.0 * m * n * n - 2.0 * n ** 3 / 3.0 + solve
    n = r.n
    return 2.0 * n * n if r.reuse else n ** 3 / 3.0 + 2.0 * n * n
