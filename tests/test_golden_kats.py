"""The reference's known-answer cases (tests/golden/reference_kats.json).

CPU: the oracle reproduces every fixture (LU / QR / Cholesky restatements,
the GMRES restatement's failure on the restart-2 case).  GPU: the solvers
reproduce them through lx.linear_solve.
"""

import json
import pathlib

import numpy as np
import pytest

from oracle import lineax_oracle as O

KATS = json.loads((pathlib.Path(__file__).resolve().parent / "golden" / "reference_kats.json").read_text())


def test_oracle_reproduces_dense_kats():
    sq = KATS["square_2x2"]
    for kind in ("lu", "qr"):
        x, res = O.solve_dense(kind, np.array(sq["matrix"]), np.array(sq["vector"]))
        assert res == 0
        np.testing.assert_allclose(x, sq["solution"], rtol=1e-14, atol=1e-14)
    for name in ("wide_2x3", "tall_3x2"):
        k = KATS[name]
        x, res = O.solve_dense("qr", np.array(k["matrix"]), np.array(k["vector"]))
        np.testing.assert_allclose(x, k["solution"], rtol=1e-12, atol=1e-12)
    tri = KATS["lower_triangular_2x2"]
    a = np.array(tri["matrix"])
    x, _ = O.solve_dense("lu", a, np.array(tri["vector"]))
    np.testing.assert_allclose(x, tri["solution"], rtol=1e-14)
    d = KATS["diagonal"]
    np.testing.assert_allclose(np.array(d["vector"]) / np.array(d["diagonal"]), d["solution"], rtol=1e-15)


def test_oracle_gmres_restart2_kat_fails():
    k = KATS["gmres_restart2_not_successful"]
    a = np.array(k["matrix"])
    b = a @ np.array(k["true_x"])
    _, res, _ = O.gmres(lambda v: a @ v, b, k["rtol"], k["atol"], restart=k["restart"])
    assert res != 0


@pytest.mark.gpu
def test_solvers_reproduce_kats(device):
    import torch

    import lineax_amd as lx

    t = lambda x: torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)
    sq = KATS["square_2x2"]
    for solver in (lx.LU(), lx.QR(), lx.AutoLinearSolver(well_posed=True)):
        x = lx.linear_solve(lx.MatrixLinearOperator(t(sq["matrix"])), t(sq["vector"]), solver).value
        np.testing.assert_allclose(x.cpu().numpy(), sq["solution"], rtol=1e-14, atol=1e-14)
    for name in ("wide_2x3", "tall_3x2"):
        k = KATS[name]
        for solver in (lx.QR(), lx.AutoLinearSolver(well_posed=None)):
            x = lx.linear_solve(lx.MatrixLinearOperator(t(k["matrix"])), t(k["vector"]), solver).value
            np.testing.assert_allclose(x.cpu().numpy(), k["solution"], rtol=1e-12, atol=1e-12)
    tri = KATS["lower_triangular_2x2"]
    op = lx.MatrixLinearOperator(t(tri["matrix"]), lx.lower_triangular_tag)
    x = lx.linear_solve(op, t(tri["vector"]), lx.Triangular()).value
    np.testing.assert_allclose(x.cpu().numpy(), tri["solution"], rtol=1e-14)
    d = KATS["diagonal"]
    op = lx.MatrixLinearOperator(torch.diag(t(d["diagonal"])), lx.diagonal_tag)
    for solver in (lx.Diagonal(), lx.AutoLinearSolver(well_posed=True)):
        x = lx.linear_solve(op, t(d["vector"]), solver).value
        np.testing.assert_allclose(x.cpu().numpy(), d["solution"], rtol=1e-15)
    k = KATS["gmres_restart2_not_successful"]
    a = np.array(k["matrix"])
    sol = lx.linear_solve(lx.MatrixLinearOperator(t(a)), t(a @ np.array(k["true_x"])),
                          lx.GMRES(rtol=k["rtol"], atol=k["atol"], restart=k["restart"]), throw=False)
    assert int(sol.result) != lx.RESULTS.successful
