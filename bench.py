"""Benchmark of the solver hot path on one MI355X.

Default workload: the uneven-batch Cholesky configuration — 20,000 SPD
float32 systems, n ~ U{8..256}, A = M M^T / n + 0.1 I, b ~ N(0, 1), seed 1 —
solved as one step per pass over the whole batch, inputs resident in HBM.

  chosen form: one persistent launch per tile class (three: orders
               129..256, 65..128, <= 64, from the autotune table), FP32
               factorisation on the matrix cores with the blocks in
               registers, FP64 refinement to normwise backward error
               <= 1e-12 (float64 x).
  plain form:  the kernels lx.Cholesky runs (factor + solve launch pair) on
               one order bucket at a time (249 buckets), FP32, no
               refinement (float32 x).  The per-bucket schedule on the
               chosen form's kernel is reported beside it.

``value`` is the chosen form's solves/s.  Per-step latency percentiles, the
plain form's figures, accuracy against the float64 oracle (on a sample of
systems), the roofline of the dominant kernel (timed with HIP events on the
launch stream) and a CPU baseline (LAPACK spotrf + spotrs per system on the
host cores, bounded sample) are reported on the same JSON line, and, under
"configs", every other benchmark configuration with its own roofline and
CPU baseline (``--configs none`` skips them).

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
    host = []
    for start, end in events:
        start.record(stream)
        h0 = time.perf_counter()
        fn()
        host.append((time.perf_counter() - h0) * 1e3)
        end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _time_steps.last_host_ms = host
    return [s.elapsed_time(e) for s, e in events], wall


def _traffic_from_profiles(kernel_prefix):
    """HBM bytes per launch from committed rocprofv3 PMC summaries, if present."""
    path = ROOT / "profiles" / "uneven_cholesky_pmc.txt"
    if not path.exists():
        return None
    m = re.search(r"^bytes_per_step:\s*([0-9.eE+]+)", path.read_text(), re.MULTILINE)
    if m is None:
        m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", path.read_text(), re.MULTILINE)
    return float(m.group(1)) if m else None


# ---------------------------------------------------------------------------
# CPU baselines: LAPACK (through scipy), the routines lineax reaches on the
# JAX CPU backend, in the precision lineax computes in for float32 data.
# ---------------------------------------------------------------------------

_POOL_TASKS = None


def _pool_chol(span):
    """Worker: LAPACK spotrf + spotrs on the systems span[0]:span[1] of the
    uneven batch left in _POOL_TASKS before the fork (one BLAS thread)."""
    import scipy.linalg.lapack as L
    from threadpoolctl import threadpool_limits

    host, order = _POOL_TASKS
    with threadpool_limits(1):
        for k in range(*span):
            i = int(order[k])
            c, info = L.spotrf(host.matrix(i), lower=0, clean=0)
            L.spotrs(c, host.rhs(i), lower=0)
    return span[1] - span[0]


def _cpu_workers():
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))


def cpu_baseline(host, seconds_target=15.0):
    """LAPACK spotrf + spotrs per system (what lx.Cholesky runs on the JAX CPU
    backend for float32 data) over a bounded random sample of the same
    systems, in single-threaded worker processes, one per host core."""
    import multiprocessing as mp

    global _POOL_TASKS
    workers = _cpu_workers()
    rng = np.random.default_rng(7)
    order = rng.permutation(host.count)
    _POOL_TASKS = (host, order)
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        # calibrate on 64 systems per worker, then size the sample to ~seconds_target
        calib = min(host.count, 64 * workers)
        chunks = [(i, min(calib, i + 16)) for i in range(0, calib, 16)]
        t0 = time.perf_counter()
        pool.map(_pool_chol, chunks)
        per = (time.perf_counter() - t0) / calib
        count = int(min(host.count, max(calib, seconds_target / max(per, 1e-9))))
        chunks = [(i, min(count, i + 32)) for i in range(0, count, 32)]
        t0 = time.perf_counter()
        done = sum(pool.map(_pool_chol, chunks))
        dt = time.perf_counter() - t0
    _POOL_TASKS = None
    return {
        "value": done / dt,
        "unit": "solves/s",
        "cores": workers,
        "kind": "port",
        "sample": f"{done} of the {host.count} systems (random, seed 7): LAPACK spotrf+spotrs per system "
                  f"(scipy, float32), {workers} single-threaded worker processes",
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
    ap.add_argument("--workload", default="uneven_cholesky",
                    choices=["uneven_cholesky", "small_lu", "large_lu", "krylov_cg", "krylov_gmres", "serving",
                             "krylov_bicgstab", "lsmr", "svd", "triangular", "normal"],
                    help="default: the uneven-batch Cholesky configuration (the headline metric)")
    ap.add_argument("--n", type=int, default=16384, help="large_lu: order (8192 or 16384)")
    ap.add_argument("--precision", default="bf16x3", choices=["bf16", "fp16", "bf16x3", "fp32"],
                    help="large_lu: Schur products")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: arrivals per second")
    ap.add_argument("--requests", type=int, default=400, help="serving: stream length")
    ap.add_argument("--refine-steps", type=int, default=2, help="serving: FP64 refinement steps (chosen form)")
    ap.add_argument("--serving-forms", default=None, help="serving: comma-separated forms to run (default: all)")
    ap.add_argument("--max-iters", type=int, default=None, help="krylov_cg: cap on CG iterations (default: lineax's)")
    ap.add_argument("--configs", default="all", choices=["all", "none"],
                    help="with the default workload: also run every other configuration and carry it under "
                         "'configs' (default all)")
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
            "by cost over GPUs with no exchange)",
            file=sys.stderr,
        )
        sys.exit(2)

    legs = {
        "uneven_cholesky": bench_uneven_cholesky,
        "small_lu": bench_small_lu,
        "large_lu": bench_large_lu,
        "krylov_cg": bench_krylov_cg,
        "krylov_gmres": bench_krylov_gmres,
        "serving": bench_serving,
        "krylov_bicgstab": bench_krylov_bicgstab,
        "lsmr": bench_lsmr,
        "svd": bench_svd,
        "triangular": bench_triangular,
        "normal": bench_normal,
    }
    out = legs[args.workload](args)
    if args.workload == "uneven_cholesky" and args.configs == "all":
        out["configs"] = run_configs(args)
    print(json.dumps(out))


# The other benchmark configurations, run after the headline by the default
# invocation and carried in its JSON line under "configs": each with its own
# roofline and CPU baseline.
CONFIG_LEGS = [
    ("small_lu_1024x32", "small_lu", {}),
    ("large_lu_8192_bf16x3", "large_lu", {"n": 8192, "precision": "bf16x3"}),
    ("large_lu_8192_bf16", "large_lu", {"n": 8192, "precision": "bf16"}),
    ("large_lu_8192_fp16", "large_lu", {"n": 8192, "precision": "fp16"}),
    ("large_lu_16384_bf16x3", "large_lu", {"n": 16384, "precision": "bf16x3"}),
    ("large_lu_16384_bf16", "large_lu", {"n": 16384, "precision": "bf16"}),
    ("large_lu_16384_fp16", "large_lu", {"n": 16384, "precision": "fp16"}),
    ("krylov_cg_2048x2048_16rhs", "krylov_cg", {}),
    ("krylov_gmres30_1024x1024_16rhs", "krylov_gmres", {"steps": 30}),
    ("serving_2k_per_s", "serving", {"rate": 2000.0}),
    ("serving_10k_per_s", "serving", {"rate": 10000.0}),
    ("serving_50k_per_s", "serving", {"rate": 50000.0}),
    # a rate below the loop's capacity, arrivals over 5 s: the latency of a
    # stable queue rather than of a backlog
    ("serving_sustained_600_per_s", "serving", {"rate": 600.0, "requests": 3000, "serving_forms": "chosen"}),
    # the configuration's 2k/s for 5 s: the rate the loop serves under that
    # load and the latency of its queue
    ("serving_sustained_2k_per_s", "serving", {"rate": 2000.0, "requests": 10000, "serving_forms": "chosen"}),
]

_SUMMARY_KEYS = ("metric", "value", "unit", "ms_per_step", "p50_ms", "p99_ms", "dtype", "config", "roofline",
                 "cpu_baseline", "eta_max", "eta_met", "factor_ms", "refine_ms", "refine_stats", "plain_form",
                 "speedup_vs_plain", "speedup_vs_plain_time", "max_backward_error_sample", "iterations_max",
                 "ms_per_iteration", "results", "max_scaled_residual", "arrival_span_s", "p50_first_quarter_ms",
                 "p50_last_quarter_ms")


def run_configs(args):
    import torch

    legs = {"small_lu": bench_small_lu, "large_lu": bench_large_lu, "krylov_cg": bench_krylov_cg,
            "krylov_gmres": bench_krylov_gmres, "serving": bench_serving}
    out = {}
    for name, leg, overrides in CONFIG_LEGS:
        t0 = time.perf_counter()
        sub = argparse.Namespace(**{**vars(args), **overrides})
        try:
            line = legs[leg](sub)
            out[name] = {k: line[k] for k in _SUMMARY_KEYS if k in line}
            if leg == "serving":
                for form in ("chosen", "plain_form"):
                    src = line.get(form if form == "chosen" else "plain_form") or {}
                    out[name][form] = {k: src.get(k) for k in ("solves_per_s", "p50_ms", "p99_ms", "completed",
                                                                "makespan_s", "achieved_tflops", "backward_error")}
        except Exception as e:  # a failed leg is reported, not fatal to the headline
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["wall_s"] = time.perf_counter() - t0
        print(f"[bench] config {name}: {out[name].get('value')} {out[name].get('unit', '')} "
              f"({out[name]['wall_s']:.1f} s)", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    return out


def bench_uneven_cholesky(args):
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
    host_ms = float(np.mean(_time_steps.last_host_ms))
    sol = chosen["sol"]
    ok = int((sol.result == 0).sum())
    eta_dev = sol.eta.cpu().numpy()
    iters = sol.iters.cpu().numpy()
    mean_ms = float(np.mean(ms))
    print(f"[bench] chosen form: {mean_ms:.3f} ms/step, p50 {_percentile(ms, 50):.3f}, ok {ok}/{host.count}",
          file=sys.stderr, flush=True)

    plain_info = None
    fused_info = None
    refined_info = None
    if not args.no_plain:
        from lineax_amd import _kernels as K
        from oracle import lineax_oracle as O

        rng = np.random.default_rng(5)
        pidx = rng.choice(host.count, size=200, replace=False)

        def eta_sample(x_of):
            return max(O.normwise_backward_error(host.matrix(i), x_of(i), host.rhs(i)) for i in pidx)

        # Plain form: what lx.Cholesky runs under a vmap over one order bucket
        # (lxa_chol_factor + lxa_chol_solve on a [k, n, n] batch), bucket by
        # bucket; the buckets' matrices and right-hand sides are resident.
        order_asc = batch.order_asc.cpu().numpy()
        buckets = []
        for order_n, s0, e0 in batch.buckets:
            idx = order_asc[s0:e0]
            mats = np.stack([host.matrix(int(i)) for i in idx])
            rhs = np.stack([host.rhs(int(i)) for i in idx]).astype(np.float32)
            buckets.append((idx, torch.as_tensor(mats, device="cuda:0"), torch.as_tensor(rhs, device="cuda:0")))
        torch.cuda.synchronize()
        plain = {}

        def run_plain():
            plain["x"] = [K.chol_solve(K.chol_factor(a_k)[0], b_k) for _, a_k, b_k in buckets]

        pms, _ = _time_steps(run_plain, max(3, args.steps // 2), 1)
        where = {}
        for (idx, _, _), xk in zip(buckets, plain["x"]):
            xh = xk.cpu().numpy().astype(np.float64)
            for j, i in enumerate(idx):
                where[int(i)] = xh[j]
        plain_info = {
            "value": host.count / (float(np.mean(pms)) / 1e3),
            "ms_per_step": float(np.mean(pms)),
            "p50_ms": _percentile(pms, 50),
            "p99_ms": _percentile(pms, 99),
            "kernels": "lxa_chol_factor + lxa_chol_solve per order bucket (the lx.Cholesky path)",
            "launches_per_step": 2 * len(batch.buckets),
            "max_backward_error_sample": eta_sample(lambda i: where[int(i)]),
            "x_dtype": "float32",
            "achieved_gbs": batch.algorithmic_bytes / (float(np.mean(pms)) / 1e3) / 1e9,
        }
        del buckets, plain
        print(f"[bench] plain form: {plain_info['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

        # The same per-bucket schedule on the register-blocked kernel of the
        # chosen form (FP32, no refinement): separates launch structure from
        # kernel quality.
        fused = {}

        def run_fused():
            fused["sol"] = batch.cholesky_solve(form="plain", out=fused.get("sol"))

        fms, _ = _time_steps(run_fused, max(3, args.steps // 2), 1)
        fx = fused["sol"].x.cpu().numpy().astype(np.float64)
        fused_info = {
            "value": host.count / (float(np.mean(fms)) / 1e3),
            "ms_per_step": float(np.mean(fms)),
            "kernels": "chol_mfma_kernel<NB, WAVES, false, float> per order bucket (no refinement)",
            "launches_per_step": len(batch.buckets),
            "max_backward_error_sample": eta_sample(
                lambda i: fx[int(host.b_offset[i]): int(host.b_offset[i]) + int(host.n[i])]),
        }
        print(f"[bench] fused per-bucket form: {fused_info['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

        # The plain form's schedule (one launch per order bucket) at the chosen
        # form's accuracy: the chosen kernel with FP64 refinement to the same
        # tolerance, bucket by bucket.  Chosen vs this isolates the batching.
        xr = torch.empty(batch.x_len, dtype=torch.float64, device="cuda:0")
        info_r = torch.empty(batch.count, dtype=torch.int32, device="cuda:0")
        eta_r = torch.empty(batch.count, dtype=torch.float64, device="cuda:0")

        def run_refined_buckets():
            for order_n, s0, e0 in batch.buckets:
                K.chol_uneven_solve(batch.n, batch.a_offset, batch.a, batch.b_offset, batch.b, xr,
                                    batch.order_asc[s0:e0], max_n=order_n, persistent=False, max_refine=10,
                                    tol=args.tol, info=info_r, eta=eta_r)

        rms, _ = _time_steps(run_refined_buckets, max(3, args.steps // 2), 1)
        refined_info = {
            "value": host.count / (float(np.mean(rms)) / 1e3),
            "ms_per_step": float(np.mean(rms)),
            "kernels": "chol_mfma_kernel<NB, WAVES, true, double> per order bucket, FP64 refinement to tol",
            "launches_per_step": len(batch.buckets),
            "device_eta_max": float(eta_r.max()),
        }
        print(f"[bench] per-bucket refined form: {refined_info['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)

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
        "traffic_source": "profiles/uneven_cholesky_pmc.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE of one step, "
                          "copy-calibrated); not measured in this run",
        "kernel": "chol_mfma_kernel<16|8|4, 4|4|1, true, double>: one persistent launch per tile class (16x16 "
                  "blocks for orders 129..256, 8x8 for 65..128, 4x4 for <= 64) per step, plus a fill and a "
                  "prepass launch; achieved over the whole step",
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
        "host_ms_per_step": host_ms,
        "systems_ok": ok,
        "device_eta_max": float(eta_dev.max()),
        "refinement_residuals_mean": float(iters.mean()),
        "accuracy": acc,
        "plain_form": plain_info,
        "speedup_vs_plain": (value / plain_info["value"]) if plain_info else None,
        "per_bucket_fused_kernel": fused_info,
        "per_bucket_refined": refined_info,
        "speedup_vs_per_bucket_refined": (value / refined_info["value"]) if refined_info else None,
        "accuracy_note": "chosen form: float64 answers refined to device eta <= tol; plain form: float32 answers, "
                         "unrefined (max_backward_error_sample); per_bucket_refined: the plain form's launch "
                         "schedule at the chosen form's accuracy",
        "roofline": roofline,
        "cpu_baseline": cpu,
        "wall_s_timed": wall,
    }
    return out


# ---------------------------------------------------------------------------
# Other configurations (one JSON line each; not the headline line)
# ---------------------------------------------------------------------------

MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "bf16x3": 2500.0 / 3, "fp32": 157.3}


def _cpu_threads():
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))


def _line(metric, value, unit, ms, args, config, dtype, roofline, cpu, extra, higher=True):
    out = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (numpy / torch generators, fixed seeds), resident in HBM", "config": config,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    out.update(extra)
    return out


def bench_small_lu(args):
    """1024 independent 32 x 32 float32 LU solves: one fused factor + solve launch
    (lxa_lu_small_solve), as linear_solve with lx.LU() runs them."""
    import torch

    from lineax_amd import _kernels as K
    from lineax_amd import workloads as W
    from oracle import lineax_oracle as O

    a_h, b_h = W.small_lu_batch(1024, 32, seed=0)
    a = torch.as_tensor(a_h, device="cuda:0")
    b = torch.as_tensor(b_h, device="cuda:0")
    res = {}

    def step():
        # what linear_solve(MatrixLinearOperator(A), b, LU()) runs on a batch of
        # small operators: one fused getrf + getrs launch that also writes the
        # factors and pivots (the solver state)
        res["x"], _, _, _ = K.lu_small_solve(a, b, want_factors=True)

    ms, _ = _time_steps(step, args.steps, args.warmup)
    t = float(np.mean(ms))
    x = res["x"].cpu().numpy().astype(np.float64)
    worst = max(O.normwise_backward_error(a_h[i].astype(np.float64), x[i], b_h[i].astype(np.float64))
                for i in range(0, 1024, 8))
    per_bytes = 4 * (32 * 32 + 2 * 32)
    achieved = 1024 * per_bytes / (t / 1e3) / 1e9
    cpu_s = _lapack_small_lu_seconds(a_h, b_h)
    return _line("solves/s", 1024 / (t / 1e3), "solves/s", t, args,
                 {"workload": "small_lu", "systems": 1024, "n": 32, "seed": 0}, "f32",
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                  "kernel": "lu_small_fused_kernel<float, 32> (getrf + getrs, factors and pivots written)",
                  "algorithmic_bytes_per_launch": 1024 * per_bytes},
                 {"value": 1024 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                  "sample": "all 1024 systems: LAPACK sgetrf+sgetrs per system (scipy, float32), one thread"},
                 {"max_backward_error_sample": worst})


_LARGE_LU_CACHE = {}


def _lapack_small_lu_seconds(a_h, b_h):
    """Seconds for LAPACK sgetrf + sgetrs over every system, one thread."""
    import scipy.linalg.lapack as L
    from threadpoolctl import threadpool_limits

    with threadpool_limits(1):
        t0 = time.perf_counter()
        for i in range(a_h.shape[0]):
            lu, piv, _ = L.sgetrf(a_h[i])
            L.sgetrs(lu, piv, b_h[i])
        return time.perf_counter() - t0


def _lapack_large_lu(n, seed=2):
    """LAPACK sgetrf + sgetrs (4 right-hand sides) on one float32 N(0,1) matrix of
    order n, BLAS on every host core: what lx.LU runs on the JAX CPU backend.
    Measured once per order (cached)."""
    if n in _LARGE_LU_CACHE:
        return _LARGE_LU_CACHE[n]
    import scipy.linalg.lapack as L
    from threadpoolctl import threadpool_limits

    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n), dtype=np.float32)
    b = rng.standard_normal((n, 4), dtype=np.float32)
    cores = _cpu_workers()
    with threadpool_limits(cores):
        t0 = time.perf_counter()
        lu, piv, info = L.sgetrf(a, overwrite_a=1)
        L.sgetrs(lu, piv, b)
        dt = time.perf_counter() - t0
    del a, lu
    res = {"value": 1.0 / dt, "unit": "systems/s (4 right-hand sides each)", "cores": cores, "kind": "port",
           "sample": f"one LAPACK sgetrf + sgetrs at n = {n} (scipy, float32, {cores} BLAS threads), "
                     f"{dt:.2f} s measured"}
    _LARGE_LU_CACHE[n] = res
    return res


def bench_large_lu(args):
    """One float32 system of order n with 4 right-hand sides: butterfly + windowed-pivot
    blocked LU (Schur products in --precision) + FP64 GMRES refinement to 1e-12 or
    10 residuals.  bf16 and fp16 factors need longer GMRES cycles (restart 30)."""
    import torch

    from lineax_amd.large_lu import LargeLU, factor_flops

    n = int(args.n)
    gen = torch.Generator(device="cpu").manual_seed(2)
    a = torch.randn((n, n), generator=gen, dtype=torch.float32).to("cuda:0")
    b = torch.randn((4, n), generator=gen, dtype=torch.float32).to("cuda:0")
    info = {}
    # bf16 / fp16 factors need longer Arnoldi cycles; GMRES's max_steps counts
    # restart cycles (lineax's dummy first pass included), so each correction
    # is bounded at 60 preconditioned iterations for them
    low = args.precision in ("bf16", "fp16")
    restart, inner_max = (30, 3) if low else (4, 50)

    def factor():
        info["fac"] = LargeLU(a, precision=args.precision)

    def refine():
        info["x"], info["eta"], info["stats"] = info["fac"].solve(b, tol=args.tol, max_refine=10, restart=restart,
                                                                 inner_max_steps=inner_max)

    steps = max(1, min(args.steps, 3))
    fms, _ = _time_steps(factor, steps, 1)
    # one untimed refinement first (the first call of a process loads its
    # kernels), except for the second-long bf16 / fp16 refinements
    sms, _ = _time_steps(refine, 1 if low else steps, 0 if low else 1)
    t_f, t_s = float(np.mean(fms)), float(np.mean(sms))
    total = t_f + t_s
    flops = factor_flops(n)
    achieved = flops / (t_f / 1e3) / 1e12
    peak = MFMA_PEAK_TFLOPS[args.precision]
    # plain form: float32 Schur products, no refinement (one per order)
    key = ("plain", n)
    if key not in _LARGE_LU_CACHE:
        plain = {}

        def plain_step():
            fac = LargeLU(a, precision="fp32")
            plain["x"] = fac.apply_inverse(b)

        pms, _ = _time_steps(plain_step, max(1, min(args.steps, 2)), 0)
        fac = info["fac"]
        xp = plain["x"].double()
        rp = fac.residual(xp, b.double())
        _LARGE_LU_CACHE[key] = {"ms": float(np.mean(pms)), "eta_max": float(fac.backward_error(xp, b.double(), rp).max()),
                                "precision": "fp32", "refinement": False}
    plain_info = _LARGE_LU_CACHE[key]
    eta = float(info["eta"].max())
    refine_stats = info["stats"]
    del info
    torch.cuda.empty_cache()
    cpu = None if args.no_cpu_baseline else _lapack_large_lu(n)
    return _line("solves/s", 1.0 / (total / 1e3), "systems/s (4 right-hand sides each)", total, args,
                 {"workload": "large_lu", "n": n, "nrhs": 4, "precision": args.precision, "seed": 2,
                  "tol_backward_error": args.tol, "max_refine": 10, "gmres_restart": restart},
                 f"f32 storage, {args.precision} Schur products, f64 refinement",
                 {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                  "traffic": None, "kernel": "factorisation (gemm_kernel Schur updates dominate)",
                  "algorithmic_flops": flops},
                 cpu,
                 {"factor_ms": t_f, "refine_ms": t_s, "eta_max": eta, "eta_met": eta <= args.tol,
                  "refine_stats": refine_stats, "plain_form": plain_info, "speedup_vs_plain_time": plain_info["ms"] / total})


def _krylov_cpu_baseline(kind, coeffs, b0, iters_gpu, rows):
    """The oracle's FP64 CG / GMRES on one right-hand side, one thread, for a bounded
    block of iterations; the per-iteration time is scaled to the GPU run's count."""
    from oracle import lineax_oracle as O

    shape = b0.shape
    mv = lambda v: O.stencil_apply(v.reshape(shape), coeffs).ravel()
    b0 = b0.ravel()
    t0 = time.perf_counter()
    if kind == "cg":
        block = 40
        O.cg(mv, b0, 1e-6, 1e-6, max_steps=block)
        per = (time.perf_counter() - t0) / block
        value = 1.0 / (per * max(iters_gpu, 1))
        unit, sample = "right-hand sides solved per second", f"{block} CG iterations on one {shape[0]}^2 grid"
    else:
        block = 2
        O.gmres(mv, b0, 1e-6, 1e-6, max_steps=block + 1, restart=30)
        per = (time.perf_counter() - t0) / block
        value = 1.0 / per
        unit, sample = "restart cycles x right-hand sides per second", f"{block} GMRES(30) cycles on one " \
                                                                         f"{shape[0]}^2 grid"
    return {"value": value, "unit": unit, "cores": 1, "kind": "port",
            "sample": sample + " (oracle, float64, one thread), time scaled to the measured iteration count",
            "seconds_per_iteration": per}


def _krylov_leg(args, kind):
    """The chosen form of the Krylov configurations: the iterate y and the
    right-hand side b in float64 (exact upcasts of float32 data), float32
    working vectors, FP64 inner products and norms, graph-captured iteration
    blocks.  lineax's own float32 CG cannot run this problem: its breakdown
    guard rcond = 2 eps(float32) n is 1.0 at n = 2048^2 (cg.py:131,173-178).

    CG is timed to convergence.  GMRES(30) is timed over `--steps` restart
    cycles: the unpreconditioned 1024^2 convection-diffusion problem needs
    thousands of cycles, so its figure is per cycle."""
    import torch

    import lineax_amd as lx
    from lineax_amd.stencil import convection_diffusion2d, poisson2d

    dev = "cuda:0"
    rows = 16
    gen = torch.Generator(device="cpu").manual_seed(3 if kind == "cg" else 4)
    if kind == "cg":
        nx = 2048
        op = poisson2d(nx, dtype=torch.float64, device=dev)
        # fused 2-sweep schedule: 9 float32 passes per point per iteration
        bytes_per_iter = 36 * nx * nx * rows

        def make(max_steps):
            return lx.CG(rtol=1e-6, atol=1e-6, max_steps=max_steps, working_dtype=torch.float32, graph=True)

        timed = make(args.max_iters)
        warm = make(70)
    else:
        nx = 1024
        op = convection_diffusion2d(nx, dtype=torch.float64, device=dev)
        m = 30
        bytes_per_iter = 4 * nx * nx * (m * m + 5 * m + 5) * rows  # per restart cycle

        def make(cycles):
            return lx.GMRES(rtol=1e-6, atol=1e-6, restart=m, max_steps=cycles + 1, working_dtype=torch.float32,
                            graph=True)

        timed = make(int(args.steps))
        warm = make(3)
    b32 = torch.randn((rows, nx, nx), generator=gen, dtype=torch.float32)
    b = b32.to(torch.float64).to(dev)
    lx.linear_solve(op, b, warm, throw=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = lx.linear_solve(op, b, timed, throw=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps = sol.stats["num_steps"].cpu().numpy()
    results = sol.result.cpu().numpy()
    # CG: iterations; GMRES: restart cycles after lineax's dummy first pass
    iters = float(steps.max()) if kind == "cg" else float(steps.max()) - 1.0
    per_iter_ms = wall * 1e3 / max(iters, 1.0)
    achieved = bytes_per_iter / (per_iter_ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        coeffs = op.coeffs
        cpu = _krylov_cpu_baseline(kind, coeffs, b32[0].double().numpy(), iters, rows)
    extra = {"iterations_max": int(steps.max()), "results": results.tolist(), "ms_per_iteration": per_iter_ms,
             "wall_s": wall}
    if kind == "cg":
        value, unit = rows / wall, "right-hand sides solved per second"
        # residual check of the returned iterate, in float64 on the device
        y = sol.value
        r = b - op.mv(y)
        scale = 1e-6 + 1e-6 * torch.abs(b)
        extra["max_scaled_residual"] = float(torch.max(torch.abs(r) / scale))
    else:
        value, unit = iters * rows / wall, "restart cycles x right-hand sides per second"
    return _line("solves/s" if kind == "cg" else "restart cycles/s", value, unit, wall * 1e3, args,
                 {"workload": f"krylov_{kind}", "grid": [nx, nx], "rhs": rows, "rtol": 1e-6, "atol": 1e-6,
                  "form": "chosen (float64 iterate, float32 working vectors)"},
                 "f32 working vectors, f64 iterate / inner products",
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                  "kernel": "one iteration (CG) / restart cycle (GMRES), all launches",
                  "algorithmic_bytes_per_iteration": bytes_per_iter},
                 cpu, extra)


def bench_krylov_cg(args):
    return _krylov_leg(args, "cg")


def bench_krylov_gmres(args):
    return _krylov_leg(args, "gmres")


def _serving_errors(reqs, ops):
    """FP64 backward errors of every completed request, on the device.
    Cholesky: eta = |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) (Rigal-Gaches).
    Least squares: eta_LS = |A^T (b - A x)|_inf / (|A|_1 (|A|_inf |x|_inf + |b|_inf)),
    the normal-equations residual, which is ~ m u_f64 at the exact solution."""
    import torch

    errs = {"qr": [], "chol": []}
    for r in reqs:
        if r.done is None:
            continue
        a = (r.a if r.a is not None else ops[r.op_id]).double()
        x, b = r.x.double(), r.b.double()
        res = b - a @ x
        den = a.abs().sum(1).max() * x.abs().max() + b.abs().max()
        if r.kind == "qr":
            errs["qr"].append((a.T @ res).abs().max() / (a.abs().sum(0).max() * den))
        else:
            errs["chol"].append(res.abs().max() / den)
    out = {}
    for k, v in errs.items():
        if v:
            t = torch.stack(v).cpu().numpy()
            out[k] = {"max": float(t.max()), "p50": float(np.median(t)), "count": int(t.size)}
    return out


def _serving_cpu_baseline(reqs, ops, count=48):
    """LAPACK per request (what lx.QR / lx.Cholesky run on the JAX CPU backend for
    float32 data): sgeqrf + sormqr (Q^T b) + strtrs for least squares, spotrf +
    spotrs for Cholesky; a reuse request solves against its operator's stored
    factorisation.  The first `count` requests of the stream, in order, BLAS on
    every host core."""
    import scipy.linalg.lapack as L
    from threadpoolctl import threadpool_limits

    sample = reqs[:count]
    host_ops = {}
    items = []
    for r in sample:
        if r.op_id not in host_ops:
            host_ops[r.op_id] = (r.a if r.a is not None else ops[r.op_id]).cpu().numpy().astype(np.float32)
        items.append((r.kind, r.op_id, r.reuse, r.b.cpu().numpy().astype(np.float32)))
    cores = _cpu_workers()
    factors = {}
    with threadpool_limits(cores):
        t0 = time.perf_counter()
        for kind, op_id, reuse, b in items:
            if kind == "qr":
                if op_id not in factors or not reuse:
                    qr, tau, _, _ = L.sgeqrf(host_ops[op_id])
                    factors[op_id] = (qr, tau)
                qr, tau = factors[op_id]
                n = qr.shape[1]
                qtb, _, _ = L.sormqr("L", "T", qr, tau, b[:, None], max(1, n))
                L.strtrs(qr[:n, :n], qtb[:n, 0], lower=0)
            else:
                if op_id not in factors or not reuse:
                    factors[op_id] = L.spotrf(host_ops[op_id], lower=0, clean=0)[0]
                L.spotrs(factors[op_id], b, lower=0)
        dt = time.perf_counter() - t0
    return {"value": len(items) / dt, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"the first {len(items)} requests of the stream, in order: LAPACK sgeqrf+sormqr+strtrs "
                      f"(least squares) / spotrf+spotrs (Cholesky), float32 (scipy, {cores} BLAS threads)"}


_SERVING_WARM = []


def bench_serving(args):
    """Poisson request stream (50% least-squares QR, 50% SPD Cholesky, 10% reuse)."""
    import torch

    from lineax_amd import serving as S

    out = {}
    cpu = None
    if not _SERVING_WARM:
        # the first requests of a process load every kernel of the loop: warm
        # both forms up on a short stream of their own before timing
        for form in ("chosen", "plain", "fifo"):
            wr, wo = S.make_request_stream(args.rate, 12, seed=99, device="cuda:0", qr_m=(64, 1024),
                                           chol_n=(16, 512))
            S.SolveServer("cuda:0", form=form, hbm_budget=4 << 30, refine_steps=args.refine_steps).run(wr, wo)
        torch.cuda.synchronize()
        _SERVING_WARM.append(True)
    forms = ("chosen", "plain", "fifo") if args.rate == 2000.0 else ("chosen", "plain")
    if getattr(args, "serving_forms", None):
        forms = tuple(args.serving_forms.split(","))
    for form in forms:
        reqs, ops = S.make_request_stream(args.rate, args.requests, seed=5, device="cuda:0")
        if cpu is None and not args.no_cpu_baseline:
            cpu = _serving_cpu_baseline(reqs, ops)
        server = S.SolveServer("cuda:0", form=form, hbm_budget=32 << 30, resident_budget=8 << 30,
                               refine_steps=args.refine_steps)
        stats = server.run(reqs, ops, time_limit=300.0)
        flops = sum(S.request_flops(r) for r in reqs if r.done is not None)
        stats["achieved_tflops"] = flops / stats["makespan_s"] / 1e12
        stats["backward_error"] = _serving_errors(reqs, ops)
        if form == "chosen":
            stats["refine_steps"] = args.refine_steps
            # a stable queue keeps the latency of late arrivals near that of
            # early ones; a backlog makes it grow through the stream
            done = sorted((r for r in reqs if r.done is not None), key=lambda r: r.arrival)
            q = max(1, len(done) // 4)
            lat = [1e3 * (r.done - r.arrival) for r in done]
            stats["p50_first_quarter_ms"] = float(np.median(lat[:q])) if lat else None
            stats["p50_last_quarter_ms"] = float(np.median(lat[-q:])) if lat else None
            stats["arrival_span_s"] = float(max(r.arrival for r in reqs))
        out[form] = stats
        del reqs, ops, server
        torch.cuda.empty_cache()
    c, p = out["chosen"], out.get("plain") or {}
    errs = [e["max"] for e in (c.get("backward_error") or {}).values()]
    eta_max = max(errs) if errs else None
    return _line("solves/s + p50/p99 latency at equal FP64-oracle error, chosen vs plain form", c["solves_per_s"],
                 "solves/s", c["makespan_s"] * 1e3, args,
                 {"workload": "serving", "rate_per_s": args.rate, "requests": args.requests, "seed": 5,
                  "mix": "50% QR m~U{64..4096} n~U{m/4..m}, 50% Cholesky n~U{16..1024}, 10% reuse"},
                 "f32 factors + f64 refinement",
                 {"bound": "mfma", "achieved": c["achieved_tflops"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                  "frac": c["achieved_tflops"] / FP32_PEAK_TFLOPS, "traffic": None,
                  "kernel": "whole request loop (QR panel + trailing kernels dominate)"},
                 cpu,
                 {"p50_ms": c["p50_ms"], "p99_ms": c["p99_ms"], "chosen": c, "plain_form": p,
                  "plain_form_kind": "static batching: each batch of arrived requests grouped by kind and size "
                                     "bucket (QR by m in 256-row buckets, Cholesky by n in 64s), FP32 kernels "
                                     "once per bucket, no refinement, no residency",
                  "fifo_form": out.get("fifo"),
                  "arrival_span_s": c.get("arrival_span_s"),
                  "p50_first_quarter_ms": c.get("p50_first_quarter_ms"),
                  "p50_last_quarter_ms": c.get("p50_last_quarter_ms"),
                  # every refined answer against the float64 data (eta_LS for least squares)
                  "eta_max": eta_max, "eta_met": eta_max is not None and eta_max <= 1e-12,
                  "speedup_vs_plain": c["solves_per_s"] / p["solves_per_s"] if p.get("solves_per_s") else None})


def _blas_threads():
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))


def bench_krylov_bicgstab(args):
    """BiCGStab on the GMRES operator (1024^2 upwind convection-diffusion, 16
    right-hand sides): float64 iterate and b, float32 working vectors, FP64
    inner products, graph-captured blocks, timed over a fixed number of
    iterations (unpreconditioned it needs thousands to reach 1e-6)."""
    import torch

    import lineax_amd as lx
    from lineax_amd.stencil import convection_diffusion2d
    from oracle import lineax_oracle as O

    dev, nx, rows = "cuda:0", 1024, 16
    iters = args.max_iters or 300
    op = convection_diffusion2d(nx, dtype=torch.float64, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(4)
    b32 = torch.randn((rows, nx, nx), generator=gen, dtype=torch.float32)
    b = b32.to(torch.float64).to(dev)

    def make(steps):
        return lx.BiCGStab(rtol=1e-6, atol=1e-6, max_steps=steps, working_dtype=torch.float32, graph=True)

    lx.linear_solve(op, b, make(40), throw=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = lx.linear_solve(op, b, make(iters), throw=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps = int(sol.stats["num_steps"].max())
    per_ms = wall * 1e3 / steps
    # 23 vector passes per iteration, float32-equivalent (the direction
    # update 4, two stencil applications 2 + 2, <r0, v> 2, s 3, <t, s> and
    # <t, t> 2, the update of y and r with the next rho and both norms 8)
    bytes_per_iter = 23 * 4 * nx * nx * rows
    achieved = bytes_per_iter / (per_ms / 1e3) / 1e9
    # residual of the returned iterate, float64 on the device, against |b|
    r = b - op.mv(sol.value)
    rel = float((torch.linalg.vector_norm(r.reshape(rows, -1), dim=1)
                 / torch.linalg.vector_norm(b.reshape(rows, -1), dim=1)).max())
    cpu = None
    if not args.no_cpu_baseline:
        coeffs = op.coeffs
        b0 = b32[0].double().numpy()
        mv = lambda v: O.stencil_apply(v.reshape(nx, nx), coeffs).ravel()
        block = 10
        c0 = time.perf_counter()
        O.bicgstab(mv, b0.ravel(), 1e-6, 1e-6, max_steps=block)
        cpu_per = (time.perf_counter() - c0) / block
        cpu = {"value": 1.0 / cpu_per, "unit": "iterations x RHS / s", "cores": 1, "kind": "port",
               "sample": f"{block} BiCGStab iterations on one {nx}^2 grid (oracle, float64, one thread)"}
    return _line("iterations x right-hand sides per second", steps * rows / wall, "iterations x RHS / s",
                 wall * 1e3, args,
                 {"workload": "krylov_bicgstab", "grid": [nx, nx], "rhs": rows, "iterations": steps,
                  "rtol": 1e-6, "atol": 1e-6, "seed": 4},
                 "f32 working vectors, f64 iterate / inner products",
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                  "kernel": "one BiCGStab iteration, all launches (bcg_* + stencil5 + copies)",
                  "algorithmic_bytes_per_iteration": bytes_per_iter},
                 cpu,
                 {"ms_per_iteration": per_ms, "max_relative_residual": rel,
                  "results": sol.result.cpu().numpy().tolist()})


def bench_lsmr(args):
    """LSMR on one dense float32 least-squares problem, 32768 x 4096: N(0, 1/n)
    entries with the columns scaled by 10^(-2 j / n) (singular values over
    about two decades), to rtol = atol = 1e-6, graph-captured iteration
    blocks.  A v and A^T u are the GEMV kernel on the same storage (no
    transposed copy); they are the dominant kernels and are timed on their
    own with HIP events for the roofline."""
    import torch

    import lineax_amd as lx
    from lineax_amd import _kernels as K
    from oracle import lineax_oracle as O

    dev, m, n = "cuda:0", 32768, 4096
    gen = torch.Generator(device="cpu").manual_seed(6)
    scale = torch.logspace(0.0, -2.0, n, dtype=torch.float64).to(torch.float32)
    a_h = torch.randn((m, n), generator=gen, dtype=torch.float32) * (scale / float(np.sqrt(n)))
    b_h = torch.randn((m,), generator=gen, dtype=torch.float32)
    a, b = a_h.to(dev), b_h.to(dev)
    op = lx.MatrixLinearOperator(a)
    solver = lx.LSMR(1e-6, 1e-6, graph=True)
    res = {}

    def step():
        res["sol"] = lx.linear_solve(op, b, solver, throw=False)

    ms, _ = _time_steps(step, args.steps, args.warmup)
    t = float(np.mean(ms))
    sol = res["sol"]
    iters = int(sol.stats["num_steps"])
    # per iteration: A v and A^T u read A (8mn); the vector passes of the
    # GEMVs and the three LSMR kernels (u: 5m, v: 5n, x/h/hbar: 7n; 4 B each)
    bytes_per_iter = 8 * m * n + 4 * (5 * m + 13 * n)
    iter_gbs = bytes_per_iter * iters / (t / 1e3) / 1e9
    # the two products alone, HIP events on the launch stream
    v = torch.randn((n,), device=dev)
    u = torch.randn((m,), device=dev)
    kms, _ = _time_steps(lambda: (K.gemv(a, v), K.gemv(a.mT, u)), 20, 3)
    kms = float(np.mean(kms))
    achieved = 8 * m * n / (kms / 1e3) / 1e9
    # least-squares optimality of the answer, float64 on the device
    a64 = a.double()
    x64 = sol.value.double()
    r64 = b.double() - a64 @ x64
    opt = float(torch.linalg.vector_norm(a64.T @ r64) / (torch.linalg.norm(a64) * torch.linalg.vector_norm(r64)))
    del a64
    cpu = None
    threads = _blas_threads()
    if not args.no_cpu_baseline:
        from threadpoolctl import threadpool_limits

        a_c, b_c = a_h.double().numpy(), b_h.double().numpy()
        block = 10
        with threadpool_limits(limits=threads):
            c0 = time.perf_counter()
            O.lsmr(lambda w: a_c @ w, lambda z: a_c.T @ z, b_c, n, 1e-6, 1e-6, max_steps=block)
            per = (time.perf_counter() - c0) / block
        cpu = {"value": 1.0 / (per * iters), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"{block} iterations of the oracle's LSMR (float64 numpy, BLAS threads) on the same "
                         f"problem, time scaled to the {iters} iterations of the GPU solve"}
    return _line("solves/s", 1e3 / t, "solves/s", t, args,
                 {"workload": "lsmr", "m": m, "n": n, "rtol": 1e-6, "atol": 1e-6, "seed": 6,
                  "column_scaling": "10^(-2 j / n)"}, "f32 (f64 scalars)",
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                  "kernel": "gemv_n4_kernel (A v) + gemv_t_part_kernel / gemv_t_reduce_kernel (A^T u), one read "
                            "of A each",
                  "ms_per_pair": kms, "algorithmic_bytes_per_pair": 8 * m * n},
                 cpu,
                 {"iterations": iters, "result": int(sol.result), "istop": int(sol.stats["istop"]),
                  "ms_per_iteration": t / max(iters, 1), "iteration_gbs": iter_gbs,
                  "iteration_frac": iter_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_iteration": bytes_per_iter,
                  "ls_optimality_AtR_over_A_R": opt})


def bench_svd(args):
    """SVD solves of 512 float32 least-squares problems of 96 x 64 (one-sided
    Jacobi SVD, one workgroup per matrix, then the pseudoinverse product)."""
    import torch

    import lineax_amd as lx
    from lineax_amd import _kernels as K
    from oracle import lineax_oracle as O

    dev, batch, m, n = "cuda:0", 512, 96, 64
    rng = np.random.default_rng(8)
    a_h = (rng.standard_normal((batch, m, n)) / np.sqrt(n)).astype(np.float32)
    b_h = rng.standard_normal((batch, m)).astype(np.float32)
    a, b = torch.as_tensor(a_h, device=dev), torch.as_tensor(b_h, device=dev)
    op = lx.MatrixLinearOperator(a)
    res = {}

    def step():
        res["sol"] = lx.linear_solve(op, b, lx.SVD(), throw=False)

    ms, _ = _time_steps(step, args.steps, args.warmup)
    t = float(np.mean(ms))
    sweeps = K.svd(a)[3].double().cpu().numpy()
    # per sweep n(n-1)/2 pairs, each 3 inner products over m (6m), the
    # rotation of two G columns (6m) and two V columns (6n); then the solve
    flops = float(np.sum(sweeps * (n * (n - 1) / 2) * (12 * m + 6 * n))) + batch * (4 * m * n + 2 * n * n)
    tflops = flops / (t / 1e3) / 1e12
    from threadpoolctl import threadpool_limits

    sample = 64
    with threadpool_limits(limits=1):
        c0 = time.perf_counter()
        refs = [O.svd_solve(a_h[i].astype(np.float64), b_h[i].astype(np.float64))[0] for i in range(sample)]
        cpu_s = time.perf_counter() - c0
    x = res["sol"].value.double().cpu().numpy()
    err = max(O.relative_error(x[i], refs[i]) for i in range(sample))
    return _line("solves/s", batch / (t / 1e3), "solves/s", t, args,
                 {"workload": "svd", "batch": batch, "m": m, "n": n, "seed": 8}, "f32 (f64 inner products)",
                 {"bound": "mfma", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                  "frac": tflops / FP32_PEAK_TFLOPS, "traffic": None,
                  "kernel": "svd_jacobi_kernel<float> (vector ALU; one workgroup per matrix, latency-bound)",
                  "algorithmic_flops_per_step": flops, "mean_sweeps": float(sweeps.mean())},
                 {"value": sample / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                  "sample": f"{sample} of the {batch} problems, float64 LAPACK SVD + pseudoinverse, one thread"},
                 {"max_rel_err_vs_fp64_oracle": err})


def bench_triangular(args):
    """256 upper-triangular float32 solves of order 2048 (Triangular solver,
    one workgroup per system, blocked substitution)."""
    import torch

    import lineax_amd as lx
    from oracle import lineax_oracle as O

    dev, batch, n = "cuda:0", 256, 2048
    gen = torch.Generator(device=dev).manual_seed(9)
    a = torch.triu(torch.randn((batch, n, n), generator=gen, dtype=torch.float32, device=dev), 1)
    a *= 0.5 / float(np.sqrt(n))
    a += torch.diag_embed(1.0 + torch.rand((batch, n), generator=gen, dtype=torch.float32, device=dev))
    b = torch.randn((batch, n), generator=gen, dtype=torch.float32, device=dev)
    op = lx.MatrixLinearOperator(a, lx.upper_triangular_tag)
    res = {}

    def step():
        res["sol"] = lx.linear_solve(op, b, lx.Triangular(), throw=False)

    ms, _ = _time_steps(step, args.steps, args.warmup)
    t = float(np.mean(ms))
    # the upper triangle once (2 n^2 B) and the vector in and out
    per = 2 * n * n + 8 * n
    achieved = batch * per / (t / 1e3) / 1e9
    import scipy.linalg as sl
    from threadpoolctl import threadpool_limits

    sample = 32
    u64 = a[:sample].double().cpu().numpy()
    b64 = b[:sample].double().cpu().numpy()
    with threadpool_limits(limits=1):
        c0 = time.perf_counter()
        refs = [sl.solve_triangular(u64[i], b64[i], lower=False) for i in range(sample)]
        cpu_s = time.perf_counter() - c0
    x = res["sol"].value[:sample].double().cpu().numpy()
    err = max(O.relative_error(x[i], refs[i]) for i in range(sample))
    return _line("solves/s", batch / (t / 1e3), "solves/s", t, args,
                 {"workload": "triangular", "batch": batch, "n": n, "seed": 9}, "f32",
                 {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "trsv_kernel (step)",
                  "algorithmic_bytes_per_launch": batch * per},
                 {"value": sample / cpu_s, "unit": "solves/s", "cores": 1, "kind": "reference",
                  "sample": f"{sample} of the {batch} systems, float64 LAPACK trtrs (scipy solve_triangular), "
                            "one thread"},
                 {"max_rel_err_vs_fp64_lapack": err})


def bench_normal(args):
    """Normal(Cholesky()) on 64 float32 least-squares problems of 2048 x 256:
    A^T A as one batched device GEMM, the Cholesky kernels, A^T b and the
    two triangular solves."""
    import torch

    import lineax_amd as lx
    from oracle import lineax_oracle as O

    dev, batch, m, n = "cuda:0", 64, 2048, 256
    rng = np.random.default_rng(10)
    a_h = (rng.standard_normal((batch, m, n)) / np.sqrt(n)).astype(np.float32)
    b_h = rng.standard_normal((batch, m)).astype(np.float32)
    a, b = torch.as_tensor(a_h, device=dev), torch.as_tensor(b_h, device=dev)
    op = lx.MatrixLinearOperator(a)
    res = {}

    def step():
        res["sol"] = lx.linear_solve(op, b, lx.Normal(lx.Cholesky()), throw=False)

    ms, _ = _time_steps(step, args.steps, args.warmup)
    t = float(np.mean(ms))
    flops = batch * (2.0 * m * n * n + n ** 3 / 3.0 + 2.0 * m * n + 2.0 * n * n)
    tflops = flops / (t / 1e3) / 1e12
    from threadpoolctl import threadpool_limits

    sample = 16
    with threadpool_limits(limits=1):
        c0 = time.perf_counter()
        refs = [O.normal_solve(a_h[i].astype(np.float64), b_h[i].astype(np.float64)) for i in range(sample)]
        cpu_s = time.perf_counter() - c0
    x = res["sol"].value.double().cpu().numpy()
    err = max(O.relative_error(x[i], refs[i]) for i in range(sample))
    return _line("solves/s", batch / (t / 1e3), "solves/s", t, args,
                 {"workload": "normal", "batch": batch, "m": m, "n": n, "seed": 10}, "f32",
                 {"bound": "mfma", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                  "frac": tflops / FP32_PEAK_TFLOPS, "traffic": None,
                  "kernel": "step: Gram GEMM (library) + chol_factor/chol_solve kernels + gemv",
                  "algorithmic_flops_per_step": flops},
                 {"value": sample / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                  "sample": f"{sample} of the {batch} problems, float64 normal equations (oracle), one thread"},
                 {"max_rel_err_vs_fp64_oracle": err})

if __name__ == "__main__":
    main()
