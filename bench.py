"""Benchmark of the solver hot path on one MI355X.

Default workload: the uneven-batch Cholesky configuration — 20,000 SPD
float32 systems, n ~ U{8..256}, A = M M^T / n + 0.1 I, b ~ N(0, 1), seed 1 —
solved as one step per pass over the whole batch, inputs resident in HBM.

  chosen form: one persistent launch, FP32 factorisation in LDS, FP64
               refinement to normwise backward error <= 1e-12 (float64 x).
  plain form:  one FP32 launch per order bucket (249 launches), no
               refinement (float32 x).

``value`` is the chosen form's solves/s.  Per-step latency percentiles, the
plain form's figures, accuracy against the float64 oracle (on a sample of
systems), the roofline of the dominant kernel (timed with HIP events on the
launch stream) and a CPU baseline (the oracle's float64 Cholesky on host
cores, bounded sample) are reported on the same JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def _percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def _time_steps(fn, steps, warmup):
    """Runs fn warmup+steps times; returns per-step milliseconds (HIP events on
    torch's current stream, which is where every kernel is launched)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for start, end in events:
        start.record(stream)
        fn()
        end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return [s.elapsed_time(e) for s, e in events], wall


def _traffic_from_profiles(kernel_prefix):
    """HBM bytes per launch from committed rocprofv3 PMC summaries, if present."""
    path = ROOT / "profiles" / "uneven_cholesky_pmc.txt"
    if not path.exists():
        return None
    m = re.search(r"bytes_per_launch:\s*([0-9.eE+]+)", path.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(host, seconds_target=15.0):
    """The oracle's float64 Cholesky (potrf + potrs, as lineax on JAX-CPU)
    over a bounded random sample of the same systems, on host threads."""
    from oracle import lineax_oracle as O

    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    rng = np.random.default_rng(7)
    order = rng.permutation(host.count)
    # calibrate on 200 systems, then size the sample to ~seconds_target
    calib = host.subset(order[:200])
    t0 = time.perf_counter()
    O.chol_uneven(calib.n, calib.a_offset, calib.a, calib.b_offset, calib.b, threads=threads)
    dt = time.perf_counter() - t0
    per = dt / 200
    count = int(min(host.count, max(200, seconds_target / max(per, 1e-9))))
    sample = host.subset(order[:count])
    t0 = time.perf_counter()
    O.chol_uneven(sample.n, sample.a_offset, sample.a, sample.b_offset, sample.b, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": count / dt,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{count} of the {host.count} systems (random, seed 7), float64 potrf+potrs per system",
    }


def accuracy_sample(host, sol, count=300):
    from oracle import lineax_oracle as O

    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(host.count, size=min(count, host.count), replace=False))
    sub = host.subset(idx)
    x_ref, _ = O.chol_uneven(sub.n, sub.a_offset, sub.a, sub.b_offset, sub.b, threads=8)
    x = sol.x.cpu().numpy().astype(np.float64)
    worst_rel, worst_eta = 0.0, 0.0
    for k, i in enumerate(idx):
        o, n = int(host.b_offset[i]), int(host.n[i])
        xi = x[o: o + n]
        ref = x_ref[int(sub.b_offset[k]): int(sub.b_offset[k]) + n]
        worst_rel = max(worst_rel, O.relative_error(xi, ref))
        worst_eta = max(worst_eta, O.normwise_backward_error(host.matrix(i), xi, host.rhs(i)))
    return {"systems_checked": int(len(idx)), "max_rel_err_vs_fp64_oracle": worst_rel, "max_backward_error": worst_eta}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 512 MiB device-to-device copy 5 times (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.ones(128 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": src.numel() * 4, "launches": 5}))
        return

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU; --gpus {args.gpus} is not run here (the uneven batch would shard "
            "by cost over GPUs with no exchange; see DESIGN.md)",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    from lineax_amd import workloads as W
    from lineax_amd.batch import UnevenBatch

    t_gen = time.perf_counter()
    host = W.uneven_spd_batch(count=args.count, seed=1)
    print(f"[bench] generated {host.count} systems in {time.perf_counter() - t_gen:.1f}s digest={host.digest()}",
          file=sys.stderr, flush=True)
    batch = UnevenBatch.from_host(host, "cuda:0")
    torch.cuda.synchronize()

    out = {}
    chosen = {}

    def run_chosen():
        chosen["sol"] = batch.cholesky_solve(form="chosen", tol=args.tol, out=chosen.get("sol"))

    ms, wall = _time_steps(run_chosen, args.steps, args.warmup)
    sol = chosen["sol"]
    ok = int((sol.result == 0).sum())
    eta_dev = sol.eta.cpu().numpy()
    iters = sol.iters.cpu().numpy()
    mean_ms = float(np.mean(ms))
    print(f"[bench] chosen form: {mean_ms:.3f} ms/step, p50 {_percentile(ms, 50):.3f}, ok {ok}/{host.count}",
          file=sys.stderr, flush=True)

    plain_info = None
    if not args.no_plain:
        plain = {}

        def run_plain():
            plain["sol"] = batch.cholesky_solve(form="plain", out=plain.get("sol"))

        pms, _ = _time_steps(run_plain, max(3, args.steps // 2), 1)
        psol = plain["sol"]
        from oracle import lineax_oracle as O

        rng = np.random.default_rng(5)
        pidx = rng.choice(host.count, size=200, replace=False)
        px = psol.x.cpu().numpy().astype(np.float64)
        peta = max(
            O.normwise_backward_error(host.matrix(i), px[int(host.b_offset[i]): int(host.b_offset[i]) + int(host.n[i])],
                                      host.rhs(i))
            for i in pidx
        )
        plain_info = {
            "value": host.count / (float(np.mean(pms)) / 1e3),
            "ms_per_step": float(np.mean(pms)),
            "p50_ms": _percentile(pms, 50),
            "p99_ms": _percentile(pms, 99),
            "launches_per_step": len(batch.buckets),
            "max_backward_error_sample": peta,
            "x_dtype": "float32",
        }
        print(f"[bench] plain form: {plain_info['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    acc = accuracy_sample(host, sol)
    value = host.count / (mean_ms / 1e3)
    achieved = batch.algorithmic_bytes / (mean_ms / 1e3) / 1e9
    traffic = _traffic_from_profiles("chol_uneven_kernel")
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "kernel": "chol_uneven_kernel<true,double> (persistent)",
        "algorithmic_bytes_per_launch": batch.algorithmic_bytes,
        "fp32_flop_view": {
            "achieved_tflops": batch.algorithmic_flops / (mean_ms / 1e3) / 1e12,
            "peak_tflops": FP32_PEAK_TFLOPS,
        },
    }
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(host)
        print(f"[bench] cpu baseline: {cpu['value']:.1f} solves/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    out = {
        "metric": "solves/s + p50/p99 latency at equal FP64-oracle error, chosen vs plain form",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean_ms,
        "p50_ms": _percentile(ms, 50),
        "p99_ms": _percentile(ms, 99),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor + f64 refinement",
        "data": "synthetic (numpy default_rng seed 1), resident in HBM",
        "config": {
            "workload": "uneven_spd_cholesky",
            "systems": host.count,
            "n_range": [8, 256],
            "form": "chosen",
            "tol_backward_error": args.tol,
            "digest": host.digest(),
        },
        "systems_ok": ok,
        "device_eta_max": float(eta_dev.max()),
        "refinement_residuals_mean": float(iters.mean()),
        "accuracy": acc,
        "plain_form": plain_info,
        "speedup_vs_plain": (value / plain_info["value"]) if plain_info else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "wall_s_timed": wall,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
