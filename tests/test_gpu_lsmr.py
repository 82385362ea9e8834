"""GPU parity of LSMR (lineax/_solver/lsmr.py) against the oracle.

The oracle is the float64 restatement O.lsmr, itself pinned against LAPACK
least squares, the pseudoinverse and scipy's LSMR.  Float64 solves match
its solution to 1e-8 relative with the same istop and RESULTS code and a
step count within one (inner products are summed in a different order);
tall, wide, square and rank-deficient operators, batch rows that stop after
different numbers of steps, an initial guess, every stopping rule and the
RESULTS mapping of lsmr.py:344-357 are covered.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def _oracle(a, b, rtol, atol, **kw):
    return O.lsmr(lambda v: a @ v, lambda u: a.T @ u, b, a.shape[1], rtol, atol, **kw)


def _solve(a, b, solver, device, dtype=torch.float64, **kw):
    return lx.linear_solve(lx.MatrixLinearOperator(_t(a, device, dtype)), _t(b, device, dtype), solver, **kw)


@pytest.mark.parametrize("shape", [(60, 25), (25, 60), (40, 40)])
def test_lsmr_dense_vs_oracle(device, rng, shape):
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    sol = _solve(a, b, lx.LSMR(1e-10, 1e-10), device)
    x_ref, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-10, 1e-10)
    print(f"shape {shape}: steps {int(sol.stats['num_steps'])} vs {steps_ref}, istop {istop_ref}", flush=True)
    assert int(sol.result) == res_ref == O.SUCCESSFUL
    assert int(sol.stats["istop"]) == istop_ref
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-8, atol=1e-9)
    # the least-squares solution (the minimum-norm one when wide)
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.pinv(a) @ b, rtol=1e-7, atol=1e-8)


def test_lsmr_estimates(device, rng):
    a = rng.standard_normal((80, 30))
    b = rng.standard_normal(80)
    # a short run, orthogonality intact: the estimates of lsmr.py:274-312
    # match the oracle's and are the true |r| and |A^T r|
    sol = _solve(a, b, lx.LSMR(0.0, 0.0, max_steps=6), device)
    _, _, steps_ref, _, st = _oracle(a, b, 0.0, 0.0, max_steps=6, stats=True)
    assert int(sol.stats["num_steps"]) == steps_ref == 6
    for key, tol in (("norm_r", 1e-10), ("norm_Ar", 1e-8), ("norm_A", 1e-12), ("cond_A", 1e-10)):
        np.testing.assert_allclose(float(sol.stats[key]), st[key], rtol=tol, err_msg=key)
    x = sol.value.cpu().numpy()
    r = b - a @ x
    np.testing.assert_allclose(float(sol.stats["norm_r"]), np.linalg.norm(r), rtol=1e-10)
    np.testing.assert_allclose(float(sol.stats["norm_Ar"]), np.linalg.norm(a.T @ r), rtol=1e-8)
    assert float(sol.stats["norm_A"]) <= np.linalg.norm(a)
    # to convergence: past the loss of orthogonality (30 steps on 30 columns)
    # rounding differences in the inner products grow to about 1e-8 in |r|
    # and |A| (cond(A), a ratio of late rotation values, moves by percents
    # under perturbations of 1e-15 and is not compared)
    sol = _solve(a, b, lx.LSMR(1e-12, 1e-12), device)
    _, _, steps_ref, _, st = _oracle(a, b, 1e-12, 1e-12, stats=True)
    assert int(sol.stats["num_steps"]) == steps_ref
    for key in ("norm_r", "norm_A"):
        np.testing.assert_allclose(float(sol.stats[key]), st[key], rtol=1e-6, err_msg=key)
    assert float(sol.stats["norm_Ar"]) < 1e-12 + 1e-12 * float(sol.stats["norm_A"]) * float(sol.stats["norm_r"])
    x = sol.value.cpu().numpy()
    np.testing.assert_allclose(float(sol.stats["norm_r"]), np.linalg.norm(b - a @ x), rtol=1e-6)
    np.testing.assert_allclose(float(sol.stats["norm_x"]), np.linalg.norm(x), rtol=1e-12)


def test_lsmr_rank_deficient(device, rng):
    a = rng.standard_normal((30, 6)) @ rng.standard_normal((6, 12))  # rank 6
    b = rng.standard_normal(30)
    sol = _solve(a, b, lx.LSMR(1e-12, 1e-12), device)
    x_ref, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-12, 1e-12)
    assert int(sol.result) == res_ref and int(sol.stats["istop"]) == istop_ref
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-7, atol=1e-8)
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.pinv(a) @ b, rtol=1e-6, atol=1e-7)


def test_lsmr_batch_rows_stop_independently(device, rng):
    a = rng.standard_normal((50, 20))
    b = rng.standard_normal((4, 50))
    b[1] = a @ rng.standard_normal(20)  # consistent: stops on the well-posed rule
    b[2] = 0.0  # x0 = 0 solves it: no steps
    b[3] *= 1e-6
    sol = _solve(a, b, lx.LSMR(1e-9, 1e-9), device)
    for i in range(4):
        x_ref, res_ref, steps_ref, istop_ref = _oracle(a, b[i], 1e-9, 1e-9)
        print(f"row {i}: steps {int(sol.stats['num_steps'][i])} vs {steps_ref}, istop {istop_ref}", flush=True)
        assert int(sol.result[i]) == res_ref
        assert int(sol.stats["istop"][i]) == istop_ref
        assert abs(int(sol.stats["num_steps"][i]) - steps_ref) <= 1
        np.testing.assert_allclose(sol.value[i].cpu().numpy(), x_ref, rtol=1e-7, atol=1e-12)
    assert int(sol.stats["num_steps"][2]) == 0 and int(sol.stats["istop"][2]) == 1
    assert float(sol.value[2].abs().max()) == 0.0
    assert int(sol.stats["istop"][1]) == 1


def test_lsmr_initial_guess(device, rng):
    a = rng.standard_normal((45, 30))
    b = rng.standard_normal(45)
    y0 = rng.standard_normal(30)
    sol = _solve(a, b, lx.LSMR(1e-10, 1e-10), device, options={"y0": _t(y0, device)})
    x_ref, res_ref, steps_ref, _ = _oracle(a, b, 1e-10, 1e-10, x0=y0)
    assert int(sol.result) == res_ref
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-8, atol=1e-9)


def test_lsmr_max_steps_rules(device, rng):
    a = rng.standard_normal((40, 40)) + 0.1 * np.eye(40)
    b = rng.standard_normal(40)
    # max_steps given with tolerances: max_steps_reached (lsmr.py:348-353)
    sol = _solve(a, b, lx.LSMR(1e-14, 1e-14, max_steps=5), device, throw=False)
    x_ref, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-14, 1e-14, max_steps=5)
    assert (res_ref, steps_ref, istop_ref) == (O.MAX_STEPS_REACHED, 5, 4)
    assert int(sol.result) == res_ref and int(sol.stats["num_steps"]) == 5 and int(sol.stats["istop"]) == 4
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-9, atol=1e-10)
    # max_steps only, no tolerances: successful after exactly max_steps (358)
    sol = _solve(a, b, lx.LSMR(0.0, 0.0, max_steps=7), device, throw=False)
    x_ref, res_ref, steps_ref, _ = _oracle(a, b, 0.0, 0.0, max_steps=7)
    assert res_ref == O.SUCCESSFUL and steps_ref == 7
    assert int(sol.result) == res_ref and int(sol.stats["num_steps"]) == 7
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-9, atol=1e-10)


def test_lsmr_default_cap_is_singular(device, rng):
    # without max_steps the cap is 10 min(m, n) and reaching it is reported
    # as singular (lsmr.py:344-347): singular values down to 1e-12, no conlim
    # stop and tolerances the finite-precision recurrence does not reach
    n = 10
    q = np.linalg.qr(rng.standard_normal((2 * n, n)))[0]
    a = q @ np.diag(np.logspace(0, -12, n)) @ np.linalg.qr(rng.standard_normal((n, n)))[0]
    b = rng.standard_normal(2 * n)
    sol = _solve(a, b, lx.LSMR(1e-20, 1e-20, conlim=np.inf), device, throw=False)
    _, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-20, 1e-20, conlim=np.inf)
    print(f"cap: steps {int(sol.stats['num_steps'])} vs {steps_ref}, istop {istop_ref}", flush=True)
    assert (res_ref, steps_ref, istop_ref) == (O.SINGULAR, 10 * n, 4)
    assert int(sol.result) == res_ref
    assert int(sol.stats["istop"]) == 4 and int(sol.stats["num_steps"]) == 10 * n


def test_lsmr_conlim(device, rng):
    a = rng.standard_normal((30, 8)) @ np.diag(np.logspace(0, -7, 8))
    b = rng.standard_normal(30)
    sol = _solve(a, b, lx.LSMR(1e-14, 1e-14, conlim=1e3), device, throw=False)
    _, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-14, 1e-14, conlim=1e3)
    assert istop_ref == 3 and res_ref == O.CONLIM
    assert int(sol.result) == res_ref and int(sol.stats["istop"]) == 3
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1


def test_lsmr_zero_operator_column_space(device, rng):
    # A^T b = 0: alpha = 0 at the start, x0 already solves the least-squares
    # problem (istop 2, no steps)
    a = np.zeros((10, 4))
    a[:5, :] = rng.standard_normal((5, 4))
    b = np.zeros(10)
    b[5:] = rng.standard_normal(5)
    sol = _solve(a, b, lx.LSMR(1e-8, 1e-8), device)
    _, res_ref, steps_ref, istop_ref = _oracle(a, b, 1e-8, 1e-8)
    assert (istop_ref, steps_ref) == (2, 0)
    assert int(sol.stats["istop"]) == 2 and int(sol.stats["num_steps"]) == 0 and int(sol.result) == res_ref
    assert float(sol.value.abs().max()) == 0.0


def test_lsmr_float32(device, rng):
    a = rng.standard_normal((70, 30))
    b = rng.standard_normal(70)
    sol = _solve(a, b, lx.LSMR(1e-5, 1e-5), device, dtype=torch.float32)
    assert sol.value.dtype == torch.float32 and int(sol.result) == lx.RESULTS.successful
    x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ls, rtol=0, atol=1e-4 * np.abs(x_ls).max())


def test_lsmr_float32_working_vectors(device, rng):
    a = rng.standard_normal((70, 30))
    b = rng.standard_normal(70)
    sol = _solve(a, b, lx.LSMR(1e-6, 1e-6, working_dtype=torch.float32), device)
    assert sol.value.dtype == torch.float64 and int(sol.result) == lx.RESULTS.successful
    x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ls, rtol=0, atol=1e-4 * np.abs(x_ls).max())


def test_lsmr_function_operator_and_transpose(device, rng):
    a = rng.standard_normal((35, 20))
    at = _t(a, device)
    struct = lx.TensorStruct((20,), torch.float64)
    # fn takes leading batch dimensions; its transpose comes from torch.func.vjp
    op = lx.FunctionLinearOperator(lambda x: x @ at.T, struct, device=device)
    b = rng.standard_normal(35)
    sol = lx.linear_solve(op, _t(b, device), lx.LSMR(1e-10, 1e-10))
    x_ref, res_ref, steps_ref, _ = _oracle(a, b, 1e-10, 1e-10)
    assert int(sol.result) == res_ref
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-8, atol=1e-9)
    # the transposed operator: a minimum-norm solve of the wide system
    c = rng.standard_normal(20)
    sol_t = lx.linear_solve(lx.MatrixLinearOperator(at).transpose(), _t(c, device), lx.LSMR(1e-10, 1e-10))
    np.testing.assert_allclose(sol_t.value.cpu().numpy(), np.linalg.pinv(a.T) @ c, rtol=1e-7, atol=1e-8)


def test_lsmr_graph_replays_match_eager(device, rng):
    # blocks of 30 iterations captured once and replayed: the same kernels in
    # the same order as the eager loop, so the answers agree bit for bit
    # (columns scaled over a decade: about 117 steps, several blocks)
    a = rng.standard_normal((300, 120)) * np.logspace(0, -1, 120)
    b = rng.standard_normal((2, 300))
    eager = _solve(a, b, lx.LSMR(1e-10, 1e-10), device)
    graph = _solve(a, b, lx.LSMR(1e-10, 1e-10, graph=True), device)
    steps = graph.stats["num_steps"].cpu().numpy()
    print(f"graph: steps {steps.tolist()}", flush=True)
    assert steps.min() > 60
    assert torch.equal(eager.stats["num_steps"], graph.stats["num_steps"])
    assert torch.equal(eager.value, graph.value)
    assert torch.equal(eager.result, graph.result)
    x_ref, res_ref, steps_ref, _ = _oracle(a, b[0], 1e-10, 1e-10)
    assert int(graph.result[0]) == res_ref
    assert abs(int(steps[0]) - steps_ref) <= 3
    np.testing.assert_allclose(graph.value[0].cpu().numpy(), x_ref, rtol=1e-6, atol=1e-8)
