"""GPU parity of lxa_chol_uneven_solve_noalloc against lxa_chol_uneven_solve.

The no-allocation entry point runs the same kernels with the work-queue
counters in the library's own device array, so its results must equal the
workspace entry point's bit for bit.  Two calls in a row, the second with a
smaller batch, check that the counters are cleared between calls (a stale
counter would skip systems: their info / x would stay at the fill values).
"""

import numpy as np
import pytest
import torch

from lineax_amd import _kernels as K
from lineax_amd import workloads as W
from lineax_amd.batch import UnevenBatch
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _solve(batch, noalloc, max_refine=10, persistent=True):
    dev = batch.a.device
    x = torch.full((int(batch.b.numel()),), float("nan"), dtype=torch.float64, device=dev)
    info = torch.full((batch.count,), 77, dtype=torch.int32, device=dev)
    iters = torch.full((batch.count,), -5, dtype=torch.int32, device=dev)
    eta = torch.full((batch.count,), -7.0, dtype=torch.float64, device=dev)
    K.chol_uneven_solve(batch.n, batch.a_offset, batch.a, batch.b_offset, batch.b, x, batch.order_desc,
                        max_n=batch.max_n, persistent=persistent, max_refine=max_refine, tol=1e-12, info=info,
                        iters=iters, eta=eta, noalloc=noalloc)
    torch.cuda.synchronize()
    return x.cpu().numpy(), info.cpu().numpy(), iters.cpu().numpy(), eta.cpu().numpy()


@pytest.mark.parametrize("sizes", [((300, 1, 256, 21), (40, 8, 180, 22))])
def test_noalloc_matches_workspace_entry_point(device, sizes):
    for count, n_min, n_max, seed in sizes:
        host = W.uneven_spd_batch(count=count, n_min=n_min, n_max=n_max, seed=seed)
        print(f"noalloc: count={count} n in [{n_min}, {n_max}] seed={seed}", flush=True)
        batch = UnevenBatch.from_host(host, device)
        got = _solve(batch, noalloc=True)
        ref = _solve(batch, noalloc=False)
        for g, r, what in zip(got, ref, ("x", "info", "iters", "eta")):
            assert np.array_equal(g, r), f"{what} differs (count={count})"
        assert not np.isnan(got[0]).any()
        assert np.all(got[1] == 0) and np.all((got[2] >= 1) & (got[2] <= 10))
        x_ref, _ = O.chol_uneven(host.n, host.a_offset, host.a, host.b_offset, host.b, threads=4)
        for i in range(0, host.count, 7):
            o, n = int(host.b_offset[i]), int(host.n[i])
            assert O.relative_error(got[0][o: o + n], x_ref[o: o + n]) < 1e-10, (count, i, n)


def test_noalloc_plain_launch_and_unrefined(device):
    host = W.uneven_spd_batch(count=64, n_min=1, n_max=256, seed=23)
    batch = UnevenBatch.from_host(host, device)
    for refine, persistent in ((0, True), (3, True), (3, False)):
        got = _solve(batch, noalloc=True, max_refine=refine, persistent=persistent)
        ref = _solve(batch, noalloc=False, max_refine=refine, persistent=persistent)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r), (refine, persistent)
        assert np.all(got[1] == 0) and not np.isnan(got[0]).any()
