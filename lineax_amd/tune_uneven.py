"""Autotuner of the uneven-batch Cholesky launch: picks, per order bucket, the
tile class that solves a batch of that order fastest, and writes the table
the library is built with (``lineax_amd/csrc/uneven_tune.h``).

Classes (chol_mfma.hip): 4 x 4 blocks on one wave (orders <= 64), 8 x 8
blocks on four waves (orders <= 128), 16 x 16 blocks on four waves (any
order <= 256).  For each order bucket n = 8, 16, ..., 128 every class that
can hold n is timed on a batch of ``count`` SPD systems of order n
(A = M M^T / n + 0.1 I, the benchmark's construction), as the chosen form
runs them (persistent launch, FP64 refinement to 1e-12).  The table keeps
the classes contiguous in n (a persistent launch per class takes one range
of the size-sorted queue): the boundaries max4 <= max8 minimise the summed
per-system time over the buckets.

Run on the GPU, from the repository root:

    python -m lineax_amd.tune_uneven [--count 4096] [--write [--out PATH]]

``--write`` rewrites uneven_tune.h (or PATH); rebuild the library afterwards.
"""

from __future__ import annotations

import argparse
import json
import pathlib
import time

import numpy as np
import torch

from . import _kernels as K
from . import workloads as W
from .batch import UnevenBatch

TABLE = pathlib.Path(__file__).resolve().parent / "csrc" / "uneven_tune.h"
BUCKETS = list(range(8, 129, 8))


def _time_class(batch: UnevenBatch, classes, reps: int = 5) -> float:
    """Seconds per batch of the chosen form with the given class boundaries."""
    dev = batch.device
    x = torch.empty(batch.x_len, dtype=torch.float64, device=dev)
    info = torch.empty(batch.count, dtype=torch.int32, device=dev)
    ws = torch.empty(256, dtype=torch.uint8, device=dev)

    def run():
        K.chol_uneven_solve(batch.n, batch.a_offset, batch.a, batch.b_offset, batch.b, x, batch.order_desc,
                            max_n=batch.max_n, persistent=True, max_refine=10, tol=1e-12, info=info,
                            workspace=ws, sorted_order=True, classes=classes)

    run()
    torch.cuda.synchronize()
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ends[0].record()
    for e in ends[1:]:
        run()
        e.record()
    torch.cuda.synchronize()
    if not bool((info == 0).all()):
        raise RuntimeError("a tuning system failed to factor")
    return float(np.median([ends[i].elapsed_time(ends[i + 1]) for i in range(reps)])) / 1e3


def measure(count: int = 4096, seed: int = 31) -> dict:
    """Per bucket: seconds per batch of `count` systems for each eligible class."""
    out = {}
    for n in BUCKETS:
        host = W.uneven_spd_batch(count=count, n_min=n, n_max=n, seed=seed + n)
        batch = UnevenBatch.from_host(host, "cuda:0")
        row = {"16": _time_class(batch, (0, 0))}
        if n <= 128:
            row["8"] = _time_class(batch, (0, 128))
        if n <= 64:
            row["4"] = _time_class(batch, (64, 128))
        out[n] = row
        print(f"[tune] n={n}: " + ", ".join(f"{k}x{k}: {v * 1e6 / count:.2f} us/system" for k, v in row.items()),
              flush=True)
        del batch
        torch.cuda.empty_cache()
    return out


def choose(times: dict) -> tuple[int, int]:
    """Contiguous boundaries (max4, max8) minimising the summed time."""
    best = None
    ns = sorted(times)
    for m4 in [0] + [n for n in ns if n <= 64]:
        for m8 in [m4] + [n for n in ns if m4 < n <= 128]:
            total = 0.0
            for n in ns:
                cls = "4" if n <= m4 else ("8" if n <= m8 else "16")
                total += times[n][cls]
            if best is None or total < best[0]:
                best = (total, m4, m8)
    # a bucket boundary n covers orders up to the next bucket's start - 1
    _, m4, m8 = best
    step = ns[1] - ns[0] if len(ns) > 1 else 8

    def widen(m, cap):
        return 0 if m == 0 else min(cap, m + step - 1)

    return widen(m4, 64), widen(m8, 128)


def write_table(max4: int, max8: int, times: dict, device_name: str, path: pathlib.Path = TABLE) -> None:
    rows = "\n".join(
        "//   n = %3d: %s" % (n, ", ".join(f"{k}x{k} {v * 1e3:.3f} ms" for k, v in sorted(r.items(), key=lambda kv: -int(kv[0]))))
        for n, r in sorted(times.items()))
    path.write_text(f"""// Autotune table of the uneven-batch Cholesky launch (chol_mfma.hip): the
// largest order each tile class runs.  Orders up to kTunedMax4 run on 4 x 4
// blocks of one wave, up to kTunedMax8 on 8 x 8 blocks of four waves, larger
// ones on 16 x 16 blocks of four waves.  Written by lineax_amd/tune_uneven.py
// on {device_name}; measured batch times per class and order bucket:
{rows}
#pragma once

namespace lxa {{
constexpr int kTunedMax4 = {max4};
constexpr int kTunedMax8 = {max8};
}}  // namespace lxa
""")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--out", default=str(TABLE), help="where --write puts the table (default: the library's)")
    args = ap.parse_args()
    t0 = time.perf_counter()
    times = measure(args.count)
    max4, max8 = choose(times)
    built = K.uneven_tuned_classes()
    print(json.dumps({"max4": max4, "max8": max8, "library_table": built, "seconds": time.perf_counter() - t0,
                      "times_s": {str(k): v for k, v in times.items()}}))
    if args.write:
        write_table(max4, max8, times, torch.cuda.get_device_name(0), pathlib.Path(args.out))


if __name__ == "__main__":
    main()
