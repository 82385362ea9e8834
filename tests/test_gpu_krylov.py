"""GPU parity of CG and GMRES against the oracle's restatement of lineax.

Same inputs, float64: solutions agree with the oracle to the solver
tolerance (x 100), RESULTS codes agree, and the number of steps agrees to
within one (reduction order differs).  Float32 and mixed working precision
are checked against the float64 answer within the stopping tolerance.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd.stencil import convection_diffusion2d, convection_diffusion_coeffs, poisson2d
from oracle import lineax_oracle as O
from tests.helpers import poisson_1d, well_conditioned

pytestmark = pytest.mark.gpu


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


@pytest.mark.parametrize("n", [1, 2, 30, 100])
def test_cg_dense_psd(device, rng, n):
    a = well_conditioned(rng, n, kind="psd")
    b = rng.standard_normal(n)
    op = lx.MatrixLinearOperator(_t(a, device), lx.positive_semidefinite_tag)
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-10, atol=1e-10))
    x_ref, res_ref, steps_ref = O.cg(lambda v: a @ v, b, 1e-10, 1e-10)
    assert int(sol.result) == res_ref == 0
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-8, atol=1e-9)
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1


def test_cg_nsd_poisson_batched(device, rng):
    a = poisson_1d(64)
    b = rng.standard_normal((3, 64))
    op = lx.MatrixLinearOperator(_t(a, device), (lx.negative_semidefinite_tag, lx.symmetric_tag))
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-10, atol=1e-10))
    for i in range(3):
        x_ref, res_ref, steps_ref = O.cg(lambda v: a @ v, b[i], 1e-10, 1e-10, negate=True)
        np.testing.assert_allclose(sol.value[i].cpu().numpy(), x_ref, rtol=1e-7, atol=1e-8)
        assert int(sol.result[i]) == res_ref
        assert abs(int(sol.stats["num_steps"][i]) - steps_ref) <= 1


def test_cg_max_steps_rules(device, rng):
    a = -poisson_1d(100)
    b = rng.standard_normal(100)
    op = lx.MatrixLinearOperator(_t(a, device), lx.positive_semidefinite_tag)
    s = lx.linear_solve(op, _t(b, device), lx.CG(0.0, 0.0, max_steps=2), throw=False)
    assert int(s.stats["num_steps"]) == 2 and int(s.result) == 0
    s = lx.linear_solve(op, _t(b, device), lx.CG(1e-12, 1e-12, max_steps=3), throw=False)
    assert int(s.stats["num_steps"]) == 3 and int(s.result) == lx.RESULTS.max_steps_reached


def test_cg_breakdown_guard_singular(device):
    a = np.diag([1.0, 1e-30])
    op = lx.MatrixLinearOperator(_t(a, device), lx.positive_semidefinite_tag)
    s = lx.linear_solve(op, _t([0.0, 1.0], device), lx.CG(1e-12, 1e-12), throw=False)
    assert int(s.result) == lx.RESULTS.singular


def test_cg_poisson_stencil_vs_oracle(device, rng):
    n = 24
    op = poisson2d(n, device=device)
    b = rng.standard_normal((n, n))
    coeffs = (4.0, -1.0, -1.0, -1.0, -1.0)
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-8, atol=1e-8))
    x_ref, res_ref, steps_ref = O.cg(lambda v: O.stencil_apply(v.reshape(n, n), coeffs).ravel(), b.ravel(), 1e-8,
                                     1e-8)
    assert int(sol.result) == res_ref == 0
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy().ravel(), x_ref, rtol=1e-6, atol=1e-7)


def test_cg_mixed_working_precision(device, rng):
    n = 32
    op = poisson2d(n, device=device)
    b = rng.standard_normal((2, n, n))
    A = O.stencil_matrix(n, n, (4.0, -1.0, -1.0, -1.0, -1.0))
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-6, atol=1e-6, working_dtype=torch.float32))
    assert sol.value.dtype == torch.float64
    assert torch.all(sol.result == 0)
    for i in range(2):
        x = sol.value[i].cpu().numpy().ravel()
        r = b[i].ravel() - A @ x
        assert np.max(np.abs(r) / (1e-6 + 1e-6 * np.abs(b[i].ravel()))) <= 1.0


def test_gmres_dense_vs_oracle(device, rng):
    # lineax tests/test_solve.py:24-38 (100 x 100 N(0,1), restart 100)
    a = rng.standard_normal((100, 100))
    true_x = rng.standard_normal(100)
    b = a @ true_x
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device),
                          lx.GMRES(rtol=1e-10, atol=1e-10, restart=100))
    np.testing.assert_allclose(sol.value.cpu().numpy(), true_x, rtol=1e-8, atol=1e-8)
    x_ref, res_ref, steps_ref = O.gmres(lambda v: a @ v, b, 1e-10, 1e-10, restart=100)
    assert int(sol.result) == res_ref == 0
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1


def test_gmres_reference_4x4_fails(device):
    a = np.array(
        [
            [0.15892892, 0.05884365, -0.60427412, 0.1891916],
            [-1.5484863, 0.93608822, 1.94888868, 1.37069667],
            [0.62687318, -0.13996738, -0.6824359, 0.30975754],
            [-0.67428635, 1.52372255, -0.88277754, 0.69633816],
        ]
    )
    true_x = np.array([0.51383273, 1.72983427, -0.43251078, -1.11764668])
    b = a @ true_x
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device),
                          lx.GMRES(rtol=1e-10, atol=1e-10, restart=2), throw=False)
    _, res_ref, steps_ref = O.gmres(lambda v: a @ v, b, 1e-10, 1e-10, restart=2)
    assert int(sol.result) != lx.RESULTS.successful
    assert int(sol.result) == res_ref
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1


@pytest.mark.parametrize("restart", [5, 20])
def test_gmres_convdiff_stencil(device, rng, restart):
    n = 16
    op = convection_diffusion2d(n, device=device)
    coeffs = op.coeffs
    b = rng.standard_normal((2, n, n))
    sol = lx.linear_solve(op, _t(b, device), lx.GMRES(rtol=1e-8, atol=1e-8, restart=restart), throw=False)
    for i in range(2):
        x_ref, res_ref, steps_ref = O.gmres(lambda v: O.stencil_apply(v.reshape(n, n), coeffs).ravel(),
                                            b[i].ravel(), 1e-8, 1e-8, restart=restart)
        assert int(sol.result[i]) == res_ref
        assert abs(int(sol.stats["num_steps"][i]) - steps_ref) <= 1
        np.testing.assert_allclose(sol.value[i].cpu().numpy().ravel(), x_ref, rtol=1e-6, atol=1e-7)


def test_gmres_identity_three_steps(device, rng):
    b = rng.standard_normal(5)
    op = lx.MatrixLinearOperator(_t(np.eye(5), device))
    sol = lx.linear_solve(op, _t(b, device), lx.GMRES(1e-12, 1e-12))
    assert int(sol.stats["num_steps"]) == 3
    np.testing.assert_allclose(sol.value.cpu().numpy(), b, rtol=1e-14)


def test_stencil_apply_matches_oracle(device, rng):
    op = convection_diffusion2d(13, device=device)
    x = rng.standard_normal((2, 13, 13))
    y = op.mv(_t(x, device)).cpu().numpy()
    for i in range(2):
        np.testing.assert_allclose(y[i], O.stencil_apply(x[i], op.coeffs), rtol=1e-13, atol=1e-13)
    yt = op.T.mv(_t(x, device)).cpu().numpy()
    A = O.stencil_matrix(13, 13, op.coeffs)
    np.testing.assert_allclose(yt[0].ravel(), A.T @ x[0].ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("nx,ny", [(13, 13), (5, 8), (37, 64), (19, 1028), (64, 2048)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_stencil_operator_shapes(device, rng, nx, ny, dtype):
    # widths that are a multiple of 4 take the four-column kernel (row blocks
    # of 16, column strips of 1024), the others one thread per point
    from lineax_amd.stencil import StencilOperator

    coeffs = convection_diffusion_coeffs(max(nx, ny))
    op = StencilOperator(nx, ny, coeffs, dtype=dtype, device=device)
    x = rng.standard_normal((3, nx, ny))
    if dtype == torch.float32:
        x = x.astype(np.float32).astype(np.float64)
    y = op.mv(_t(x, device, dtype)).double().cpu().numpy()
    tol = 1e-13 if dtype == torch.float64 else 1e-6
    for i in range(3):
        np.testing.assert_allclose(y[i], O.stencil_apply(x[i], coeffs), rtol=tol, atol=tol)


def test_cg_grad_matches_closed_form(device, rng):
    a = well_conditioned(rng, 8, kind="psd")
    b = rng.standard_normal(8)
    g = rng.standard_normal(8)
    B = _t(b, device).requires_grad_(True)
    op = lx.MatrixLinearOperator(_t(a, device), lx.positive_semidefinite_tag)
    x = lx.linear_solve(op, B, lx.CG(rtol=1e-12, atol=1e-12)).value
    (x * _t(g, device)).sum().backward()
    np.testing.assert_allclose(B.grad.cpu().numpy(), np.linalg.solve(a.T, g), rtol=1e-8, atol=1e-9)


def test_cg_graph_replays_match_eager(device, rng):
    # graph capture of 30-iteration blocks gives the eager iterates bit for bit,
    # on every replay, with eager work between solves
    n = 48
    op = poisson2d(n, device=device)
    b = _t(rng.standard_normal((3, n, n)), device)
    eager = lx.linear_solve(op, b, lx.CG(rtol=1e-9, atol=1e-9), throw=False)
    for _ in range(3):
        g = lx.linear_solve(op, b, lx.CG(rtol=1e-9, atol=1e-9, graph=True), throw=False)
        _ = (b * 2.0).sum().item()  # unrelated eager work between the solves
        assert torch.equal(g.value, eager.value)
        assert torch.equal(g.stats["num_steps"], eager.stats["num_steps"])
        assert torch.equal(g.result, eager.result)


def test_gmres_graph_replays_match_eager(device, rng):
    n = 20
    op = convection_diffusion2d(n, device=device)
    b = _t(rng.standard_normal((2, n, n)), device)
    eager = lx.linear_solve(op, b, lx.GMRES(rtol=1e-9, atol=1e-9, restart=10), throw=False)
    for _ in range(3):
        g = lx.linear_solve(op, b, lx.GMRES(rtol=1e-9, atol=1e-9, restart=10, graph=True), throw=False)
        _ = (b * 2.0).sum().item()
        assert torch.equal(g.value, eager.value)
        assert torch.equal(g.stats["num_steps"], eager.stats["num_steps"])
        assert torch.equal(g.result, eager.result)


# --------------------------------------------------------------------------
# Fused two-sweep CG for stencil operators (lxa_cg_stencil_*)
# --------------------------------------------------------------------------

def _stencil_mv(nx, ny, coeffs):
    return lambda v: O.stencil_apply(v.reshape(nx, ny), coeffs).ravel()


@pytest.mark.parametrize("nx,ny", [(1, 1), (5, 3), (70, 300), (129, 257)])
def test_cg_fused_stencil_vs_oracle(device, rng, nx, ny):
    # tile edges in both directions (64-row, 256-column tiles), float64
    coeffs = (4.0, -1.0, -1.0, -1.0, -1.0)
    op = poisson2d(nx, ny, device=device)
    b = rng.standard_normal((2, nx, ny))
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-9, atol=1e-9))
    for i in range(2):
        x_ref, res_ref, steps_ref = O.cg(_stencil_mv(nx, ny, coeffs), b[i].ravel(), 1e-9, 1e-9)
        assert int(sol.result[i]) == res_ref == 0
        assert abs(int(sol.stats["num_steps"][i]) - steps_ref) <= 1
        np.testing.assert_allclose(sol.value[i].cpu().numpy().ravel(), x_ref, rtol=1e-6, atol=1e-7)


def test_cg_fused_matches_unfused(device, rng):
    # the fused sweeps and the six-kernel iteration run the same recurrence:
    # equal step counts, iterates equal to rounding of the reduction order
    nx, ny = 40, 260
    op = poisson2d(nx, ny, device=device)
    b = _t(rng.standard_normal((3, nx, ny)), device)
    f = lx.linear_solve(op, b, lx.CG(rtol=1e-10, atol=1e-10), throw=False)
    u = lx.linear_solve(op, b, lx.CG(rtol=1e-10, atol=1e-10, fused=False), throw=False)
    assert torch.equal(f.result, u.result)
    assert torch.max(torch.abs(f.stats["num_steps"] - u.stats["num_steps"])) <= 1
    scale = float(torch.max(torch.abs(u.value)))
    assert float(torch.max(torch.abs(f.value - u.value))) <= 1e-8 * scale


def test_cg_fused_nsd_stencil(device, rng):
    # a negative definite stencil: CG runs on -A and negates (cg.py:100-101, 224-225)
    from lineax_amd.stencil import StencilOperator

    nx, ny = 33, 47
    coeffs = (-4.0, 1.0, 1.0, 1.0, 1.0)
    op = StencilOperator(nx, ny, coeffs, device=device, tags=(lx.negative_semidefinite_tag, lx.symmetric_tag))
    b = rng.standard_normal((nx, ny))
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-10, atol=1e-10))
    x_ref, res_ref, steps_ref = O.cg(_stencil_mv(nx, ny, coeffs), b.ravel(), 1e-10, 1e-10, negate=True)
    assert int(sol.result) == res_ref == 0
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy().ravel(), x_ref, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("exact_f32_b", [True, False])
def test_cg_fused_mixed_precision(device, rng, exact_f32_b):
    # float64 iterate, float32 working vectors; b read in float32 only when exact
    nx = ny = 64
    coeffs = (4.0, -1.0, -1.0, -1.0, -1.0)
    b = rng.standard_normal((2, nx, ny))
    if exact_f32_b:
        b = b.astype(np.float32).astype(np.float64)
    sol = lx.linear_solve(poisson2d(nx, ny, device=device), _t(b, device),
                          lx.CG(rtol=1e-6, atol=1e-6, working_dtype=torch.float32))
    assert torch.all(sol.result == 0)
    A = O.stencil_matrix(nx, ny, coeffs)
    for i in range(2):
        x = sol.value[i].cpu().numpy().ravel()
        r = b[i].ravel() - A @ x
        assert np.max(np.abs(r) / (1e-6 + 1e-6 * np.abs(b[i].ravel()))) <= 1.0


def test_cg_fused_float32_vs_unfused(device, rng):
    nx = ny = 48
    b = _t(rng.standard_normal((2, nx, ny)), device, torch.float32)
    op = poisson2d(nx, ny, dtype=torch.float32, device=device)
    f = lx.linear_solve(op, b, lx.CG(rtol=1e-4, atol=1e-4), throw=False)
    u = lx.linear_solve(op, b, lx.CG(rtol=1e-4, atol=1e-4, fused=False), throw=False)
    assert torch.equal(f.result, u.result)
    assert torch.max(torch.abs(f.stats["num_steps"] - u.stats["num_steps"])) <= 2
    scale = float(torch.max(torch.abs(u.value)))
    assert float(torch.max(torch.abs(f.value - u.value))) <= 1e-3 * scale


def test_cg_fused_max_steps_rules(device, rng):
    nx, ny = 20, 30
    op = poisson2d(nx, ny, device=device)
    b = _t(rng.standard_normal((nx, ny)), device)
    s = lx.linear_solve(op, b, lx.CG(0.0, 0.0, max_steps=2), throw=False)
    assert int(s.stats["num_steps"]) == 2 and int(s.result) == 0
    s = lx.linear_solve(op, b, lx.CG(1e-14, 1e-14, max_steps=3), throw=False)
    assert int(s.stats["num_steps"]) == 3 and int(s.result) == lx.RESULTS.max_steps_reached
    # a row that starts converged (b = 0) never iterates beside one that does
    b2 = torch.stack([b, torch.zeros_like(b)])
    s = lx.linear_solve(op, b2, lx.CG(1e-8, 1e-8), throw=False)
    assert int(s.stats["num_steps"][1]) == 0 and int(s.stats["num_steps"][0]) > 0
    assert torch.all(s.value[1] == 0)


@pytest.mark.parametrize("stabilise_every,ny", [(10, 270), (7, 270), (7, 272)])
def test_cg_fused_graph_replays_match_eager(device, rng, stabilise_every, ny):
    # odd stabilisation intervals double the captured block (two direction
    # buffers); widths that are a multiple of 4 take the full-row kernels
    nx = 50
    op = poisson2d(nx, ny, device=device)
    b = _t(rng.standard_normal((2, nx, ny)), device)
    kw = dict(rtol=1e-9, atol=1e-9, stabilise_every=stabilise_every)
    eager = lx.linear_solve(op, b, lx.CG(**kw), throw=False)
    for _ in range(3):
        g = lx.linear_solve(op, b, lx.CG(graph=True, **kw), throw=False)
        _ = (b * 2.0).sum().item()
        assert torch.equal(g.value, eager.value)
        assert torch.equal(g.stats["num_steps"], eager.stats["num_steps"])
        assert torch.equal(g.result, eager.result)


# --------------------------------------------------------------------------
# Fused stencil Arnoldi steps (lxa_gmres_stencil_step)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nx,ny,restart", [(20, 20, 10), (37, 50, 7), (40, 1024, 5), (17, 3, 30)])
def test_gmres_fused_matches_unfused(device, rng, nx, ny, restart):
    # the one-sweep Arnoldi step runs lineax's classical Gram-Schmidt with the
    # basis stored unnormalised: equal RESULTS and step counts, iterates equal
    # to rounding
    from lineax_amd.stencil import StencilOperator, convection_diffusion_coeffs

    op = StencilOperator(nx, ny, convection_diffusion_coeffs(max(nx, ny)), device=device)
    b = _t(rng.standard_normal((2, nx, ny)), device)
    kw = dict(rtol=1e-9, atol=1e-9, restart=restart, max_steps=60)
    f = lx.linear_solve(op, b, lx.GMRES(**kw), throw=False)
    u = lx.linear_solve(op, b, lx.GMRES(fused=False, **kw), throw=False)
    assert torch.equal(f.result, u.result)
    assert torch.max(torch.abs(f.stats["num_steps"] - u.stats["num_steps"])) <= 1
    scale = float(torch.max(torch.abs(u.value)))
    assert float(torch.max(torch.abs(f.value - u.value))) <= 1e-7 * scale


def test_gmres_fused_vs_oracle(device, rng):
    from lineax_amd.stencil import convection_diffusion_coeffs

    n = 24
    coeffs = convection_diffusion_coeffs(n)
    op = convection_diffusion2d(n, device=device)
    b = rng.standard_normal((n, n))
    sol = lx.linear_solve(op, _t(b, device), lx.GMRES(rtol=1e-10, atol=1e-10, restart=20))
    x_ref, res_ref, steps_ref = O.gmres(lambda v: O.stencil_apply(v.reshape(n, n), coeffs).ravel(), b.ravel(), 1e-10,
                                       1e-10, restart=20)
    assert int(sol.result) == res_ref == 0
    assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
    np.testing.assert_allclose(sol.value.cpu().numpy().ravel(), x_ref, rtol=1e-7, atol=1e-8)


def test_gmres_fused_breakdown_identity_stencil(device, rng):
    # A = I as a stencil: the first Arnoldi step breaks down (w' = 0); the
    # fused and unfused paths give lineax's RESULTS and step counts
    from lineax_amd.stencil import StencilOperator

    op = StencilOperator(12, 9, (1.0, 0.0, 0.0, 0.0, 0.0), device=device)
    b = _t(rng.standard_normal((3, 12, 9)), device)
    f = lx.linear_solve(op, b, lx.GMRES(rtol=1e-8, atol=1e-8, restart=5), throw=False)
    u = lx.linear_solve(op, b, lx.GMRES(rtol=1e-8, atol=1e-8, restart=5, fused=False), throw=False)
    assert torch.equal(f.result, u.result)
    assert torch.equal(f.stats["num_steps"], u.stats["num_steps"])
    assert float(torch.max(torch.abs(f.value - b))) <= 1e-12


def test_gmres_fused_float32_working(device, rng):
    n = 64
    op = convection_diffusion2d(n, device=device)
    b = _t(rng.standard_normal((2, n, n)), device)
    kw = dict(rtol=1e-5, atol=1e-5, restart=30, working_dtype=torch.float32, max_steps=200)
    f = lx.linear_solve(op, b, lx.GMRES(**kw), throw=False)
    u = lx.linear_solve(op, b, lx.GMRES(fused=False, **kw), throw=False)
    assert torch.equal(f.result, u.result)
    assert torch.max(torch.abs(f.stats["num_steps"] - u.stats["num_steps"])) <= 2
    r = b - op.mv(f.value)
    assert float(torch.max(torch.abs(r) / (1e-5 + 1e-5 * torch.abs(b)))) <= 1.0


def test_gmres_fused_graph_replays_match_eager(device, rng):
    n = 30
    op = convection_diffusion2d(n, device=device)
    b = _t(rng.standard_normal((2, n, n)), device)
    kw = dict(rtol=1e-9, atol=1e-9, restart=10)
    eager = lx.linear_solve(op, b, lx.GMRES(**kw), throw=False)
    for _ in range(3):
        g = lx.linear_solve(op, b, lx.GMRES(graph=True, **kw), throw=False)
        _ = (b * 2.0).sum().item()
        assert torch.equal(g.value, eager.value)
        assert torch.equal(g.stats["num_steps"], eager.stats["num_steps"])
        assert torch.equal(g.result, eager.result)


# --- RESULTS rules pinned exactly against the oracle


@pytest.mark.parametrize("case", ["rotation", "shift1", "shift3"])
def test_gmres_stagnation_matches_oracle(device, case):
    # gmres.py:175-179, 228-230: cycles that do not decrease the residual norm
    # count towards stagnation_iters; these operators make every cycle's
    # step exactly zero, so the stop is deterministic.
    if case == "rotation":
        a = np.array([[0.0, 1.0], [-1.0, 0.0]])
        b = np.array([1.0, 0.0])
        restart, stag = 1, 3
    else:
        n = 6
        a = np.roll(np.eye(n), 1, axis=0)
        b = np.eye(n)[0]
        restart, stag = (1 if case == "shift1" else 3), 4
    x_ref, res_ref, steps_ref = O.gmres(lambda v: a @ v, b, 1e-12, 1e-12, restart=restart, stagnation_iters=stag)
    assert res_ref == lx.RESULTS.stagnation
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device),
                          lx.GMRES(rtol=1e-12, atol=1e-12, restart=restart, stagnation_iters=stag), throw=False)
    assert int(sol.result) == res_ref
    assert int(sol.stats["num_steps"]) == steps_ref
    np.testing.assert_array_equal(sol.value.cpu().numpy(), x_ref)


@pytest.mark.parametrize("c", [3e-13, 6e-13])
def test_cg_breakdown_guard_scale_matches_oracle(device, c):
    # 100 * rcond = 100 * 2 eps * 10 = 4.44e-13 separates the two cases
    # (lineax/_misc.py:30-38, cg.py:173-178): NaN step below, a solve above.
    n = 10
    b = np.linspace(1.0, 2.0, n)
    x_ref, res_ref, steps_ref = O.cg(lambda v: c * v, b, 1e-10, 1e-10)
    op = lx.MatrixLinearOperator(_t(c * np.eye(n), device), lx.positive_semidefinite_tag)
    sol = lx.linear_solve(op, _t(b, device), lx.CG(rtol=1e-10, atol=1e-10), throw=False)
    assert int(sol.result) == res_ref
    x = sol.value.cpu().numpy()
    if np.isnan(x_ref).any():
        assert np.isnan(x).all()
    else:
        np.testing.assert_allclose(x, x_ref, rtol=1e-12)
        assert abs(int(sol.stats["num_steps"]) - steps_ref) <= 1
