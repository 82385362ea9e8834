"""The benchmark configurations at their full sizes, on the GPU, with checks
anchored on the oracle or on size-independent properties.

* Krylov CG: the 2048 x 2048 five-point Poisson operator (n = 4,194,304),
  16 right-hand sides (seed 3), rtol = atol = 1e-6, run to convergence with
  float32 working vectors, a float64 iterate and graph-captured iteration
  blocks: every right-hand side converges (RESULTS.successful), the true
  FP64 residual of the returned iterate meets lineax's scaled test, and the
  step counts agree across right-hand sides.  Then 40 float64 CG steps on
  one right-hand side match the oracle's restatement of lineax's CG
  (lineax/_solver/cg.py:114-227) to 1e-9 relative.
* Krylov GMRES: the 1024 x 1024 upwind convection-diffusion operator
  (Peclet 10, velocity (1, 1)), restart 30, 16 right-hand sides (seed 4):
  the iterate after two restart cycles on one right-hand side matches the
  oracle's GMRES (lineax/_solver/gmres.py:106-413, max_steps = 3: lineax's
  first pass is a dummy step) to 1e-9 relative, and all 16 right-hand sides
  of the batched solve agree with their own single-RHS solves.
* Serving: the request loop over the configuration's size ranges (least
  squares m ~ U{64..4096}, n ~ U{m/4..m}; Cholesky n ~ U{16..1024}; 10%
  reuse): every FP64-refined answer reaches eta <= 1e-12 (least squares:
  the normal-equations residual eta_LS) against the float64 data.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd import serving as S
from lineax_amd.stencil import convection_diffusion2d, convection_diffusion_coeffs, poisson2d
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _rhs(seed, rows, nx):
    gen = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn((rows, nx, nx), generator=gen, dtype=torch.float32)


def test_cg_poisson_2048_sixteen_rhs_converge(device):
    nx, rows = 2048, 16
    print(f"config CG: Poisson {nx}^2 x {rows} RHS, float32 working vectors, graph", flush=True)
    op = poisson2d(nx, dtype=torch.float64, device=device)
    b = _rhs(3, rows, nx).to(torch.float64).to(device)
    solver = lx.CG(rtol=1e-6, atol=1e-6, working_dtype=torch.float32, graph=True)
    sol = lx.linear_solve(op, b, solver, throw=False)
    res = sol.result.cpu().numpy()
    steps = sol.stats["num_steps"].cpu().numpy()
    print(f"config CG: results {res.tolist()} steps {steps.min()}..{steps.max()}", flush=True)
    assert np.all(res == lx.RESULTS.successful)
    # random right-hand sides of one operator converge at nearly one rate
    assert steps.max() <= 1.05 * steps.min()
    r = b - op.mv(sol.value)
    scaled = float(torch.max(torch.abs(r) / (1e-6 + 1e-6 * torch.abs(b))))
    print(f"config CG: max scaled true residual {scaled:.3f}", flush=True)
    assert scaled <= 1.0


def test_cg_poisson_2048_forty_steps_vs_oracle(device):
    nx = 2048
    coeffs = (4.0, -1.0, -1.0, -1.0, -1.0)
    b = _rhs(3, 1, nx)[0].double().numpy()
    print("config CG: 40 float64 steps vs oracle", flush=True)
    op = poisson2d(nx, dtype=torch.float64, device=device)
    sol = lx.linear_solve(op, torch.as_tensor(b, device=device), lx.CG(rtol=1e-6, atol=1e-6, max_steps=40),
                          throw=False)
    y_ref, res_ref, steps_ref = O.cg(lambda v: O.stencil_apply(v.reshape(nx, nx), coeffs).ravel(), b.ravel(),
                                     1e-6, 1e-6, max_steps=40)
    assert steps_ref == 40 and int(sol.stats["num_steps"]) == 40
    assert int(sol.result) == res_ref == lx.RESULTS.max_steps_reached
    err = O.relative_error(sol.value.cpu().numpy().ravel(), y_ref)
    print(f"config CG: 40-step iterate relative error vs oracle {err:.2e}", flush=True)
    assert err <= 1e-9


def test_gmres_convdiff_1024_two_cycles_vs_oracle(device):
    nx, rows, m = 1024, 16, 30
    coeffs = convection_diffusion_coeffs(nx)
    op = convection_diffusion2d(nx, dtype=torch.float64, device=device)
    b32 = _rhs(4, rows, nx)
    b = b32.to(torch.float64).to(device)
    print(f"config GMRES({m}): {nx}^2 x {rows} RHS, two cycles", flush=True)
    solver = lx.GMRES(rtol=1e-6, atol=1e-6, restart=m, max_steps=3)
    sol = lx.linear_solve(op, b, solver, throw=False)
    assert torch.all(sol.result == lx.RESULTS.max_steps_reached)
    assert torch.all(sol.stats["num_steps"] == 3)
    b0 = b32[0].double().numpy().ravel()
    y_ref, res_ref, steps_ref = O.gmres(lambda v: O.stencil_apply(v.reshape(nx, nx), coeffs).ravel(), b0, 1e-6,
                                        1e-6, max_steps=3, restart=m)
    assert steps_ref == 3 and res_ref == lx.RESULTS.max_steps_reached
    err = O.relative_error(sol.value[0].cpu().numpy().ravel(), y_ref)
    print(f"config GMRES: two-cycle iterate relative error vs oracle {err:.2e}", flush=True)
    assert err <= 1e-9
    # the batched solve equals each right-hand side solved alone
    for i in (5, 15):
        one = lx.linear_solve(op, b[i], solver, throw=False)
        assert O.relative_error(sol.value[i].cpu().numpy(), one.value.cpu().numpy()) <= 1e-12


def test_gmres_convdiff_1024_float32_working_graph(device):
    """The chosen form of the GMRES configuration (float32 basis, FP64
    Hessenberg and norms, one graph-captured cycle) tracks the float64 solve
    over ten cycles to float32 accuracy."""
    nx, m = 1024, 30
    op = convection_diffusion2d(nx, dtype=torch.float64, device=device)
    b = _rhs(4, 2, nx).to(torch.float64).to(device)
    kw = dict(rtol=1e-6, atol=1e-6, restart=m, max_steps=11)
    ref = lx.linear_solve(op, b, lx.GMRES(**kw), throw=False)
    got = lx.linear_solve(op, b, lx.GMRES(working_dtype=torch.float32, graph=True, **kw), throw=False)
    assert torch.equal(got.result, ref.result)
    scale = float(torch.max(torch.abs(ref.value)))
    diff = float(torch.max(torch.abs(got.value - ref.value)))
    print(f"config GMRES float32 working: max diff {diff / scale:.2e} relative", flush=True)
    assert diff <= 1e-4 * scale


def _check_refined(r, ops):
    a = (r.a if r.a is not None else ops[r.op_id]).double().cpu().numpy()
    b = r.b.double().cpu().numpy()
    x = r.x.double().cpu().numpy()
    if r.kind == "qr":
        return O.lstsq_backward_error(a, x, b)
    return O.normwise_backward_error(a, x, b)


def test_request_loop_at_config_sizes(device):
    reqs, ops = S.make_request_stream(2000.0, 48, seed=5, device=device, qr_m=(64, 4096), chol_n=(16, 1024),
                                      reuse_fraction=0.1)
    sizes = [(r.kind, r.m, r.n) for r in reqs]
    print(f"config serving: {len(reqs)} requests, largest QR m = {max(m for k, m, n in sizes if k == 'qr')}, "
          f"largest Cholesky n = {max((n for k, m, n in sizes if k == 'chol'), default=0)}", flush=True)
    server = S.SolveServer(device, form="chosen", hbm_budget=8 << 30, resident_budget=1 << 30)
    stats = server.run(reqs, ops)
    assert stats["completed"] == len(reqs)
    worst = {"qr": 0.0, "chol": 0.0}
    for r in reqs:
        assert r.x.dtype == torch.float64
        eta = _check_refined(r, ops)
        worst[r.kind] = max(worst[r.kind], eta)
        assert eta <= 1e-12, (r.rid, r.kind, r.m, r.n, eta)
    print(f"config serving: worst eta_LS {worst['qr']:.2e}, worst eta (Cholesky) {worst['chol']:.2e}", flush=True)
