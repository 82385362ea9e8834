"""GPU parity of the uneven-batch Cholesky launch.

Chosen form (persistent launches, one per register-tile class; FP32 factor
+ FP64 refinement): each
float64 solution must match the oracle's float64 Cholesky solve of the same
(exact, upcast) data to 1e-10 relative, and have normwise backward error
<= 1e-12 both as reported by the kernel and as recomputed on the host.
Plain form (per-bucket FP32 launches): backward error < 64 n eps32.
Each result must be identical alone and in the mix.
"""

import numpy as np
import pytest
import torch

from lineax_amd import workloads as W
from lineax_amd.batch import UnevenBatch
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu

EPS32 = np.finfo(np.float32).eps


@pytest.fixture(scope="module")
def host():
    h = W.uneven_spd_batch(count=240, n_min=1, n_max=256, seed=11)
    # Force the edges into the mix.
    return h


def _x(sol, host, i):
    o = int(host.b_offset[i])
    return sol.x[o: o + int(host.n[i])].cpu().numpy().astype(np.float64)


def test_chosen_form_matches_oracle(device, host):
    batch = UnevenBatch.from_host(host, device)
    sol = batch.cholesky_solve(form="chosen", tol=1e-12)
    torch.cuda.synchronize()
    x_ref, info_ref = O.chol_uneven(host.n, host.a_offset, host.a, host.b_offset, host.b, threads=4)
    assert np.all(info_ref == 0)
    assert torch.all(sol.info == 0) and torch.all(sol.result == 0)
    eta_dev = sol.eta.cpu().numpy()
    iters = sol.iters.cpu().numpy()
    for i in range(host.count):
        a = host.matrix(i).astype(np.float64)
        b = host.rhs(i).astype(np.float64)
        xi = _x(sol, host, i)
        ref = x_ref[host.b_offset[i]: host.b_offset[i] + host.n[i]]
        assert O.relative_error(xi, ref) < 1e-10, (i, int(host.n[i]))
        assert O.normwise_backward_error(a, xi, b) <= 1e-12
        assert 0 <= eta_dev[i] <= 1e-12 and 1 <= iters[i] <= 10


def test_plain_form_backward_stable(device, host):
    batch = UnevenBatch.from_host(host, device)
    sol = batch.cholesky_solve(form="plain")
    assert sol.x.dtype == torch.float32
    assert torch.all(sol.info == 0)
    for i in range(host.count):
        n = int(host.n[i])
        eta = O.normwise_backward_error(host.matrix(i), _x(sol, host, i), host.rhs(i))
        assert eta < 64 * n * EPS32, (i, n, eta)


def test_same_alone_and_in_the_mix(device, host):
    batch = UnevenBatch.from_host(host, device)
    mixed = batch.cholesky_solve(form="chosen", tol=1e-12)
    pick = [0, 7, 100, 239]
    alone_host = host.subset(pick)
    alone = UnevenBatch.from_host(alone_host, device).cholesky_solve(form="chosen", tol=1e-12)
    for k, i in enumerate(pick):
        np.testing.assert_array_equal(_x(alone, alone_host, k), _x(mixed, host, i))


def test_failing_system_does_not_poison_others(device, host):
    idx = list(range(20))
    sub = host.subset(idx)
    bad = sub.matrix(5).copy()
    bad[0, 0] = -1.0
    o = int(sub.a_offset[5])
    n5 = int(sub.n[5])
    sub.a[o: o + n5 * n5] = bad.reshape(-1)
    sol = UnevenBatch.from_host(sub, device).cholesky_solve(form="chosen", tol=1e-12)
    res = sol.result.cpu().numpy()
    assert res[5] == 2 and np.all(np.isnan(_x(sol, sub, 5)))
    others = [i for i in idx if i != 5]
    assert np.all(res[others] == 0)
    ok = UnevenBatch.from_host(host.subset(others), device).cholesky_solve(form="chosen", tol=1e-12)
    for k, i in enumerate(others):
        np.testing.assert_array_equal(_x(ok, host.subset(others), k), _x(sol, sub, i))


def test_persistent_queue_reuse(device, host):
    # Two launches in a row reuse the queue counter; results must repeat bitwise.
    batch = UnevenBatch.from_host(host, device)
    s1 = batch.cholesky_solve(form="chosen", tol=1e-12)
    x1 = s1.x.clone()
    s2 = batch.cholesky_solve(form="chosen", tol=1e-12)
    assert torch.equal(x1, s2.x)


@pytest.mark.parametrize("orders", [
    [127, 128, 129, 130, 175, 176, 177, 178, 255, 256],  # all three tile classes, both sides of each split
    [100, 129, 150, 176, 176],                            # largest order 176: the 11x11 and 8x8 classes only
])
def test_tile_class_boundaries(device, orders):
    # Orders <= 128 run 8 x 8 register tiles, 129..176 run 11 x 11 (two
    # workgroups per compute unit, LDS sized by the class's largest order),
    # 177..256 run 16 x 16; each class is its own persistent launch, and the
    # plain form launches one class per order bucket.
    rng = np.random.default_rng(23)
    mats, rhss = [], []
    for n in orders:
        m = rng.standard_normal((n, n))
        mats.append((m @ m.T / n + 0.1 * np.eye(n)).astype(np.float32))
        rhss.append(rng.standard_normal(n).astype(np.float32))
    host = W.pack_uneven(orders, mats, rhss)
    batch = UnevenBatch.from_host(host, device)
    chosen = batch.cholesky_solve(form="chosen", tol=1e-12)
    plain = batch.cholesky_solve(form="plain")
    x_ref, info_ref = O.chol_uneven(host.n, host.a_offset, host.a, host.b_offset, host.b, threads=4)
    assert np.all(info_ref == 0) and torch.all(chosen.info == 0) and torch.all(plain.info == 0)
    for i, n in enumerate(orders):
        a = host.matrix(i).astype(np.float64)
        b = host.rhs(i).astype(np.float64)
        xi = _x(chosen, host, i)
        ref = x_ref[host.b_offset[i]: host.b_offset[i] + n]
        assert O.relative_error(xi, ref) < 1e-10, (i, n)
        assert O.normwise_backward_error(a, xi, b) <= 1e-12, (i, n)
        assert O.normwise_backward_error(a, _x(plain, host, i), b) < 64 * n * EPS32, (i, n)
