"""GPU parity of the request loop and its kernels.

* lxa_chol_blocked_factor / _solve (uneven orders up to 1024) against the
  oracle's float64 Cholesky of the same float32 data: normwise backward
  error <= 1e-5 (float32 factorisation), and RESULTS-style failure (info,
  NaN factor) on a non-positive-definite system.
* lxa_qr_factor_step with systems at different columns (one joins later)
  gives the factors of lxa_qr_factor to float32 rounding (1e-4 relative).
* The request loop: the chosen form's FP64-refined answers reach a normwise
  backward error of 1e-12 against the float64 data (least squares: the
  normal-equations residual |A^T (b - A x)|_inf / (|A|_1 (|A|_inf |x|_inf +
  |b|_inf))), ill-conditioned near-square least-squares systems included;
  each request solved alone through the loop gives the same answer as in the
  mix to 1e-10 relative (refined) — the batch can change the panel kernel's
  row split and so the float32 summation order, which refinement removes;
  without refinement the chosen form agrees with the plain form to 1e-5
  relative and both are float32-accurate; reuse requests are answered from
  the resident factorisation.
"""

import numpy as np
import pytest
import torch

from lineax_amd import _kernels as K
from lineax_amd import serving as S
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _spd(rng, n):
    m = rng.standard_normal((n, n))
    return (m @ m.T / n + 0.1 * np.eye(n)).astype(np.float32)


def test_chol_blocked_uneven_vs_oracle(device, rng):
    ns = [1, 5, 31, 32, 33, 100, 257, 513, 1024]
    mats = [_spd(rng, n) for n in ns]
    rhs = [rng.standard_normal(n).astype(np.float32) for n in ns]
    d = np.zeros(len(ns), dtype=K.QR_DESC)
    a_off = b_off = 0
    for i, n in enumerate(ns):
        d[i]["a_offset"], d[i]["b_offset"], d[i]["m"], d[i]["n"], d[i]["lda"], d[i]["ldb"] = a_off, b_off, n, n, n, n
        a_off += n * n
        b_off += n
    a = torch.as_tensor(np.concatenate([x.ravel() for x in mats]), device=device)
    b = torch.as_tensor(np.concatenate(rhs), device=device)
    info = torch.full((len(ns),), 7, dtype=torch.int32, device=device)
    dd = K.descs_to_device(d, device)
    a0 = a.clone()
    K.chol_blocked_factor_raw(dd, len(ns), a, info)
    # the strictly lower triangles are not touched
    for i, n in enumerate(ns):
        o = int(d[i]["a_offset"])
        low = torch.tril(torch.ones(n, n, dtype=torch.bool, device=device), -1)
        assert torch.equal(a[o: o + n * n].view(n, n)[low], a0[o: o + n * n].view(n, n)[low])
    K.chol_blocked_solve_raw(dd, len(ns), a, b)
    assert info.cpu().tolist() == [0] * len(ns)
    x = b.cpu().numpy().astype(np.float64)
    for i, n in enumerate(ns):
        xi = x[int(d[i]["b_offset"]): int(d[i]["b_offset"]) + n]
        assert O.normwise_backward_error(mats[i].astype(np.float64), xi, rhs[i].astype(np.float64)) <= 1e-5
        ref = O.solve_dense("cholesky", mats[i].astype(np.float64), rhs[i].astype(np.float64))[0]
        assert np.max(np.abs(xi - ref)) <= 1e-3 * np.max(np.abs(ref))


def test_chol_blocked_reports_non_pd(device, rng):
    n = 70
    a = _spd(rng, n)
    a[40, 40] = -5.0
    d = np.zeros(1, dtype=K.QR_DESC)
    d[0]["m"], d[0]["n"], d[0]["lda"], d[0]["ldb"] = n, n, n, n
    at = torch.as_tensor(a.ravel(), device=device)
    info = torch.zeros(1, dtype=torch.int32, device=device)
    K.chol_blocked_factor_raw(K.descs_to_device(d, device), 1, at, info)
    assert int(info.item()) > 0
    assert torch.isnan(at).all()


def test_qr_steps_with_staggered_columns(device, rng):
    shapes = [(700, 300), (300, 200)]
    mats = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    nb = K.qr_block(torch.float32)
    d = np.zeros(2, dtype=K.QR_DESC)
    offs = {"a": 0, "tau": 0, "t": 0, "w": 0}
    for i, (m, n) in enumerate(shapes):
        fp = S.qr_footprint(m, n, nb)
        d[i]["a_offset"], d[i]["tau_offset"], d[i]["t_offset"], d[i]["w_offset"] = (offs["a"], offs["tau"],
                                                                                     offs["t"], offs["w"])
        d[i]["m"], d[i]["n"], d[i]["lda"] = m, n, n
        for key in offs:
            offs[key] += fp[key]
    a = torch.as_tensor(np.concatenate([x.ravel() for x in mats]), device=device)
    tau = torch.zeros(offs["tau"], device=device)
    tblk = torch.zeros(offs["t"], device=device)
    work = torch.zeros(offs["w"], device=device)
    dd = K.descs_to_device(d, device)
    # system 1 joins after system 0 has done 5 panels
    k0, k1 = 0, None
    steps = 0
    while True:
        active_k = [k0 if k0 < 300 else 10**9, (k1 if k1 is not None and k1 < 200 else 10**9)]
        if active_k[0] >= 10**9 and active_k[1] >= 10**9:
            break
        kcol = torch.tensor(active_k, dtype=torch.int32).to(device)
        rows = max(shapes[i][0] - k for i, k in enumerate(active_k) if k < 10**9)
        tc = max(shapes[i][1] - k - 1 for i, k in enumerate(active_k) if k < 10**9)
        K.qr_factor_step_raw(dd, 2, kcol, rows, tc, a, tau, tblk, work)
        steps += 1
        if k0 < 300:
            k0 += nb
        if k1 is not None and k1 < 200:
            k1 += nb
        if steps == 5:
            k1 = 0
    got = a.cpu().numpy()
    for i, (m, n) in enumerate(shapes):
        ref, _, _ = K.qr_factor(torch.as_tensor(mats[i], device=device))
        gi = got[int(d[i]["a_offset"]): int(d[i]["a_offset"]) + m * n].reshape(m, n)
        r = ref.cpu().numpy()
        assert np.max(np.abs(gi - r)) <= 1e-4 * np.max(np.abs(r))


REFINED_ETA = 1e-12  # FP64-refined answers: normwise backward error bound


def _check_answer(r, ops):
    """float32 answers (plain form, or no refinement) to float32 accuracy;
    float64 (refined) answers to a backward error of 1e-12 against the float64
    data (least squares: the normal-equations residual eta_LS)."""
    a = (r.a if r.a is not None else ops[r.op_id]).double().cpu().numpy()
    b = r.b.double().cpu().numpy()
    x = r.x.double().cpu().numpy()
    refined = r.x.dtype == torch.float64
    if r.kind == "qr":
        if refined:
            assert O.lstsq_backward_error(a, x, b) <= REFINED_ETA, (r.rid, r.m, r.n)
            return
        # least-squares optimality A^T (A x - b) ~ 0, to float32 accuracy
        res = a @ x - b
        fa = np.linalg.norm(a)
        assert np.linalg.norm(a.T @ res) <= 1e-4 * fa * (fa * np.linalg.norm(x) + np.linalg.norm(b)), (r.rid, r.m, r.n)
    else:
        assert O.normwise_backward_error(a, x, b) <= (REFINED_ETA if refined else 1e-5), (r.rid, r.n)


def test_request_loop_chosen_matches_float64_and_alone(device):
    reqs, ops = S.make_request_stream(2000.0, 40, seed=11, device=device, qr_m=(64, 400), chol_n=(16, 300),
                                      reuse_fraction=0.2)
    server = S.SolveServer(device, form="chosen", hbm_budget=1 << 30, resident_budget=256 << 20)
    stats = server.run(reqs, ops)
    assert stats["completed"] == len(reqs)
    assert any(r.reuse for r in reqs)
    for r in reqs:
        assert r.x.dtype == torch.float64
        _check_answer(r, ops)
    # each request alone gives the same answer as in the mix (refined answers:
    # within 1e-10 relative)
    for r in reqs[:12]:
        alone = S.Request(rid=r.rid, kind=r.kind, arrival=0.0, m=r.m, n=r.n, op_id=r.op_id, reuse=False,
                          a=(r.a if r.a is not None else ops[r.op_id]), b=r.b)
        S.SolveServer(device, form="chosen", hbm_budget=1 << 28).run([alone], ops)
        scale = float(torch.max(torch.abs(alone.x)))
        assert float(torch.max(torch.abs(alone.x - r.x))) <= 1e-10 * max(1.0, scale), (r.rid, r.kind)


def test_request_loop_plain_matches_chosen(device):
    """Without refinement the chosen form runs the plain form's float32
    kernels in batches: the answers agree to float32 rounding."""
    reqs, ops = S.make_request_stream(5000.0, 16, seed=12, device=device, qr_m=(64, 300), chol_n=(16, 200))
    S.SolveServer(device, form="chosen", hbm_budget=1 << 29, refine_steps=0).run(reqs, ops)
    chosen = [r.x.clone() for r in reqs]
    assert all(x.dtype == torch.float32 for x in chosen)
    reqs2, ops2 = S.make_request_stream(5000.0, 16, seed=12, device=device, qr_m=(64, 300), chol_n=(16, 200))
    stats = S.SolveServer(device, form="plain", hbm_budget=1 << 29).run(reqs2, ops2)
    assert stats["completed"] == 16
    for x1, r in zip(chosen, reqs2):
        _check_answer(r, ops2)
        scale = float(torch.max(torch.abs(x1)))
        assert float(torch.max(torch.abs(x1 - r.x))) <= 1e-5 * max(1.0, scale)


def test_request_loop_refinement_near_square_least_squares(device):
    """Gaussian least-squares problems with m - n small are the ill-conditioned
    part of the mix (kappa ~ 4m/(m-n+1), heavier-tailed when square); the
    semi-normal-equations refinement reaches the FP64 bound on them."""
    gen = torch.Generator(device=device)
    reqs, ops = [], {}
    shapes = [(512, 512), (511, 510), (700, 690), (1024, 1020), (300, 75), (1500, 1490)]
    for i, (m, n) in enumerate(shapes):
        gen.manual_seed(100 + i)
        a = torch.randn((m, n), generator=gen, device=device, dtype=torch.float32)
        b = torch.randn((m,), generator=gen, device=device, dtype=torch.float32)
        reqs.append(S.Request(rid=i, kind="qr", arrival=0.0, m=m, n=n, op_id=i, reuse=False, a=a, b=b))
        ops[i] = a
    # a reuse request of the squarest operator with a new right-hand side
    gen.manual_seed(99)
    b2 = torch.randn((512,), generator=gen, device=device, dtype=torch.float32)
    reqs.append(S.Request(rid=len(shapes), kind="qr", arrival=0.02, m=512, n=512, op_id=0, reuse=True, a=None, b=b2))
    stats = S.SolveServer(device, form="chosen", hbm_budget=1 << 30, refine_steps=4).run(reqs, ops)
    assert stats["completed"] == len(reqs)
    for r in reqs:
        _check_answer(r, ops)
    # square systems: the least-squares solution solves A x = b
    for r in (reqs[0], reqs[-1]):
        a = ops[r.op_id].double().cpu().numpy()
        assert O.normwise_backward_error(a, r.x.cpu().numpy(), r.b.double().cpu().numpy()) <= REFINED_ETA


def test_request_loop_refinement_steps_zero_is_float32(device):
    reqs, ops = S.make_request_stream(5000.0, 8, seed=14, device=device, qr_m=(64, 200), chol_n=(16, 100))
    S.SolveServer(device, form="chosen", hbm_budget=1 << 28, refine_steps=0).run(reqs, ops)
    for r in reqs:
        assert r.x.dtype == torch.float32
        _check_answer(r, ops)


def test_request_loop_admission_respects_budget(device):
    reqs, ops = S.make_request_stream(50000.0, 30, seed=13, device=device, qr_m=(512, 1024), chol_n=(16, 64),
                                      reuse_fraction=0.0)
    budget = 24 << 20
    server = S.SolveServer(device, form="chosen", hbm_budget=budget, resident_budget=0)
    stats = server.run(reqs, ops)
    assert stats["completed"] == 30
    assert stats["arena_high_water_bytes"] <= budget
    for r in reqs:
        _check_answer(r, ops)


def test_static_and_fifo_plain_forms_agree(device):
    """The static-batching plain form runs the FIFO form's float32 kernels over
    batches: the answers agree to float32 rounding, and every request of an
    overloaded stream completes."""
    out = {}
    for form in ("plain", "fifo"):
        reqs, ops = S.make_request_stream(20000.0, 24, seed=15, device=device, qr_m=(64, 600), chol_n=(16, 300))
        stats = S.SolveServer(device, form=form, hbm_budget=1 << 29).run(reqs, ops)
        assert stats["completed"] == 24
        for r in reqs:
            assert r.x.dtype == torch.float32
            _check_answer(r, ops)
        out[form] = [r.x.clone() for r in reqs]
    for xs, xf in zip(out["plain"], out["fifo"]):
        scale = float(torch.max(torch.abs(xf)))
        assert float(torch.max(torch.abs(xs - xf))) <= 1e-5 * max(1.0, scale)


def test_small_cholesky_launches_on_two_streams(device):
    # The loop alternates its solve streams, so the tile-class launches of
    # consecutive ticks overlap on the GPU; each stream has its own
    # work-queue scratch (one shared scratch let a launch take the other's
    # queue entries and leave those systems unsolved).
    srv = S.SolveServer(device, form="chosen", hbm_budget=1 << 28)

    def batch(seed, count):
        g = np.random.default_rng(seed)
        reqs = []
        for i in range(count):
            n = int(g.integers(129, 257))
            m = g.standard_normal((n, n))
            a = torch.as_tensor((m @ m.T / n + 0.1 * np.eye(n)).astype(np.float32), device=device)
            b = torch.as_tensor(g.standard_normal(n).astype(np.float32), device=device)
            reqs.append(S.Request(i, "chol", 0.0, n, n, i, False, a, b))
        return reqs

    r1, r2 = batch(1, 600), batch(2, 600)
    s1, s2 = torch.cuda.Stream(device), torch.cuda.Stream(device)
    torch.cuda.synchronize(device)
    with torch.cuda.stream(s1):
        x1 = srv._chol_small_solve(r1, 0)
    with torch.cuda.stream(s2):
        x2 = srv._chol_small_solve(r2, 1)
    torch.cuda.synchronize(device)
    worst = 0.0
    for reqs, xs in ((r1, x1), (r2, x2)):
        for r, x in zip(reqs, xs):
            a, b, xv = r.a.double(), r.b.double(), x.double()
            den = a.abs().sum(1).max() * xv.abs().max() + b.abs().max()
            worst = max(worst, float(((b - a @ xv).abs().max() / den).item()))
    assert worst <= 1e-12
