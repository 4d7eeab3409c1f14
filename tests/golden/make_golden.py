"""Writes the golden fixtures under tests/golden/.

* ``known_answers.json`` -- literal input/output pairs taken from the reference's own
  tests (each entry names the reference file:line it comes from).
* ``small_projections.npz`` -- small seeded designs with their exact demeaned outputs
  z - D x, computed here with numpy (dense normal equations + pseudo-inverse, one
  refinement step).  Independent of every implementation in this repository except the
  closed form in ``oracle.oracle.exact_demean``, which the tests cross-check against a
  second closed form (scipy LSQR).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import exact_demean, sparse_design  # noqa: E402

KNOWN = {
    "design_ops": {
        "source": "crates/within/src/operator/design/tests.rs:6-57",
        "codes": [[0, 1, 1, 2, 0], [0, 0, 1, 2, 3]],
        "x": [1.0, 2.0, 3.0, 10.0, 20.0, 30.0, 40.0],
        "Dx": [11.0, 12.0, 22.0, 33.0, 41.0],
        "r": [1.0, 2.0, 3.0, 4.0, 5.0],
        "Dtr": [6.0, 5.0, 4.0, 3.0, 3.0, 4.0, 5.0],
        "n_dofs": 7,
    },
    "locality_sort": {
        "source": "crates/within/src/domain.rs:433-446",
        "codes": [[2, 0, 1, 0], [0, 0, 1, 1]],
        "obs_perm": [1, 3, 2, 0],
        "sorted": [True, False],
        "permuted_codes": [[0, 0, 1, 2], [0, 1, 1, 0]],
    },
    "nested_rescan": {
        "source": "crates/within/src/domain.rs:449-457",
        "codes": [[3, 0, 2, 1], [1, 0, 1, 0]],
        "sorted": [True, True],
    },
    "keeps_sorted": {
        "source": "crates/within/src/domain.rs:459-466",
        "codes": [[0, 0, 1, 2], [1, 0, 1, 0]],
        "obs_perm": None,
        "sorted": [True, False],
    },
    "layout": {
        "source": "crates/within/src/domain.rs:482-518",
        "levels": [[0, 0, 1, 1], [0, 2, 1, 0], [0, 0, 1, 1]],
        "ncols": [3, 1, 1],
        "offsets": [0, 6, 9],
        "n_dofs": 11,
    },
    "lsmr": {
        "source": "crates/schwarz-precond/src/lsmr/tests/solve.rs:61-97, breakdown.rs:8-104",
        "overdetermined": {"A": [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]], "b": [1, 2, 3, 3],
                           "x": [1, 2, 3], "tol": 1e-10, "diag": [0.5, 0.5, 1.0]},
        "beta_zero": {"A": [[1, 0], [0, 0]], "b": [5, 0], "x": [5, 0], "iterations": 1,
                      "stop": "ResidualTolerance", "tol": 1e-12},
        "alpha1_zero": {"A": [[1], [0]], "b": [0, 1], "x": [0], "iterations": 0,
                        "stop": "InitialNormalEquationResidualZero", "residual_norm": 1.0, "tol": 1e-12},
        "inconsistent_breakdown": {"A": [[1, 0], [0, 0]], "b": [5, 3], "x": [5, 0], "iterations": 1,
                                   "stop": "NormalEquationTolerance", "residual_norm": 3.0, "tol": 1e-12},
        "underdetermined": {"A": [[1, 0, 0], [0, 1, 0]], "b": [1, 2], "x": [1, 2, 0], "tol": 1e-12},
        "vec_norm": {"cases": [[[], 0.0], [[3.0, 4.0], 5.0]]},
    },
}


def small_cases():
    rng = np.random.default_rng(20240601)
    cases = {}
    # two factors, unweighted
    n = 400
    codes = np.column_stack([rng.integers(0, 23, n), rng.integers(0, 7, n)]).astype(np.uint32)
    y = rng.standard_normal((n, 2))
    cases["two_factor"] = (codes, y, None)
    # three factors, weighted, with an unobserved level
    n = 600
    codes = np.column_stack([rng.integers(0, 40, n), rng.integers(0, 11, n), rng.integers(0, 4, n)]).astype(np.uint32)
    codes[codes[:, 0] == 5, 0] = 6
    y = rng.standard_normal((n, 3))
    w = rng.uniform(0.5, 2.0, n)
    cases["three_factor_weighted"] = (codes, y, w)
    # disconnected pair graph (two blocks)
    n = 300
    a = rng.integers(0, 10, n)
    b = np.where(a < 5, rng.integers(0, 3, n), rng.integers(3, 6, n))
    codes = np.column_stack([a, b]).astype(np.uint32)
    y = rng.standard_normal((n, 1))
    cases["disconnected"] = (codes, y, None)
    out = {}
    for name, (codes, y, w) in cases.items():
        D = sparse_design(codes)
        out[f"{name}__codes"] = codes
        out[f"{name}__y"] = y
        if w is not None:
            out[f"{name}__w"] = w
        out[f"{name}__demeaned"] = exact_demean(D, y, w)
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(KNOWN, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "small_projections.npz"), **small_cases())
    print("wrote", os.listdir(HERE))
