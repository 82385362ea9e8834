"""Writes reference_kats.json: known-answer cases from lineax's own tests.

Inputs and expected outputs are copied as numbers from the cited test
functions; for the two non-square cases the reference test compares with
jnp.linalg.lstsq, so the expected outputs here are numpy's float64 lstsq of
the same data.  Run from the repository root:

    python tests/golden/make_reference_kats.py
"""

import json
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def main():
    kats = {}
    # lineax tests/test_solve.py:41-48 and :63-74 (LU, QR, SVD)
    kats["square_2x2"] = {
        "source": "tests/test_solve.py:41-48, 63-74",
        "matrix": [[1.0, 5.0], [-2.0, -2.0]],
        "vector": [3.0, 4.0],
        "solution": [-3.25, 1.25],
        "solvers": ["LU", "QR", "AutoLinearSolver(well_posed=True)"],
    }
    # tests/test_solve.py:50-60 (a diagonal operator over a pytree, flattened)
    kats["diagonal"] = {
        "source": "tests/test_solve.py:50-60",
        "diagonal": [8.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0],
        "vector": [4.0, 7.0, 8.0, 9.0, 2.0, 10.0, 12.0],
        "solution": [0.5, 7.0, 4.0, 3.0, 0.5, 2.0, 2.0],
        "solvers": ["Diagonal", "AutoLinearSolver(well_posed=True)"],
    }
    # tests/test_solve.py:99-108 (lower triangular)
    kats["lower_triangular_2x2"] = {
        "source": "tests/test_solve.py:99-108",
        "matrix": [[1.0, 0.0], [-2.0, -2.0]],
        "vector": [3.0, 4.0],
        "solution": [3.0, -5.0],
        "solvers": ["Triangular"],
    }
    # tests/test_singular.py:113-128 and :131-145 (compared with lstsq there)
    for name, mat, vec, src in (
        ("wide_2x3", [[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]], [3.0, 4.0], "tests/test_singular.py:113-128"),
        ("tall_3x2", [[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]], [3.0, 4.0, 5.0], "tests/test_singular.py:131-145"),
    ):
        sol = np.linalg.lstsq(np.array(mat), np.array(vec), rcond=None)[0]
        kats[name] = {"source": src, "matrix": mat, "vector": vec, "solution": sol.tolist(),
                      "solvers": ["QR", "AutoLinearSolver(well_posed=None)"]}
    # tests/test_singular.py:75-99: GMRES(restart=2) must not report success
    kats["gmres_restart2_not_successful"] = {
        "source": "tests/test_singular.py:75-99",
        "matrix": [
            [0.15892892, 0.05884365, -0.60427412, 0.1891916],
            [-1.5484863, 0.93608822, 1.94888868, 1.37069667],
            [0.62687318, -0.13996738, -0.6824359, 0.30975754],
            [-0.67428635, 1.52372255, -0.88277754, 0.69633816],
        ],
        "true_x": [0.51383273, 1.72983427, -0.43251078, -1.11764668],
        "rtol": 1e-10,
        "atol": 1e-10,
        "restart": 2,
        "expect": "result != successful",
    }
    (HERE / "reference_kats.json").write_text(json.dumps(kats, indent=1) + "\n")


if __name__ == "__main__":
    main()
