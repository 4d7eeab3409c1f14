"""Many independent problems in one call (solve_many): every problem's output must equal the
same Solver's own solve_batch, whatever launch shape runs them (one stream without graphs,
several streams with captured iteration graphs of 1-4 iterations, the autotuner's choice).
The per-run computation is the same sequence of kernels in every shape; the extra iterations
a group or the lag-one poll runs past convergence are gated off, so the comparison is tight."""
import numpy as np
import pytest

import within_amd as wa
from tests._configs import cfg4_problems

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def problems():
    out = []
    for codes, Y in cfg4_problems(seed=41, n_problems=20, lo=3.3, hi=5.0):
        out.append((wa.Solver(codes), Y))
    rng = np.random.default_rng(7)
    # weighted, and one problem wider than one device batch (two chunks)
    n = 30_000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 1500, n), rng.integers(0, 90, n)]).astype(np.uint32))
    out.append((wa.Solver(codes, weights=np.exp(rng.normal(0, 0.5, n))), np.asfortranarray(rng.standard_normal((n, 70)))))
    # a slope term (host back-transform of x) under the Schwarz preconditioner
    n = 20_000
    a, b = rng.integers(0, 800, n).astype(np.uint32), rng.integers(0, 60, n).astype(np.uint32)
    z = rng.standard_normal(n)
    out.append((wa.Solver([wa.Effect(a, True), wa.Effect(b, True, [z])]), np.asfortranarray(rng.standard_normal((n, 3)))))
    return out


@pytest.fixture(scope="module")
def singles(problems):
    opts = wa.LsmrOptions(tol=1e-8)
    return [s.solve_batch(Y, opts) for s, Y in problems]


@pytest.mark.parametrize("shape", [(1, 1, False), (4, 1, True), (8, 2, True), (16, 4, True)])
def test_many_matches_individual_solves(gpu, problems, singles, shape):
    wa.set_many_shape(*shape)
    try:
        res = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], wa.LsmrOptions(tol=1e-8))
    finally:
        wa.set_many_shape(0)
    got = wa.last_many_shape()
    assert (got["streams"], got["iterations_per_graph"], got["graph"]) == shape
    worst_d = worst_x = 0.0
    for p, (r, ref) in enumerate(zip(res, singles)):
        assert r.converged == ref.converged, p
        assert r.iterations == ref.iterations, p
        assert all(r.converged), p
        worst_d = max(worst_d, max(rel(r.demeaned[:, j], ref.demeaned[:, j]) for j in range(r.demeaned.shape[1])))
        worst_x = max(worst_x, max(rel(r.x[:, j], ref.x[:, j]) for j in range(r.x.shape[1])))
        np.testing.assert_allclose(r.true_residual, ref.true_residual, rtol=1e-9, atol=1e-15)
    print(f"shape {shape}: demeaned {worst_d:.2e}, x {worst_x:.2e} vs solve_batch over {len(res)} problems")
    assert worst_d <= 1e-13 and worst_x <= 1e-12


def test_many_autotuned_mixed_precision(gpu, problems):
    opts = wa.LsmrOptions(tol=1e-8, precision="mixed")
    res = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], opts)
    shape = wa.last_many_shape()
    # a second call reuses the cached shape
    res2 = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], opts)
    assert wa.last_many_shape() == shape
    for p, ((s, Y), r, r2) in enumerate(zip(problems, res, res2)):
        ref = s.solve_batch(Y, opts)
        assert all(r.converged), p
        assert r.iterations == ref.iterations, p
        for j in range(Y.shape[1]):
            assert rel(r.demeaned[:, j], ref.demeaned[:, j]) <= 1e-13
            assert rel(r2.demeaned[:, j], ref.demeaned[:, j]) <= 1e-13
    print(f"autotuned shape {shape}")


def test_many_maxiter_and_zero_rhs(gpu, problems):
    s, Y = problems[0]
    Z = np.zeros_like(Y)
    opts = wa.LsmrOptions(tol=1e-14, maxiter=5)
    wa.set_many_shape(4, 2, True)
    try:
        r_cap, r_zero = wa.solve_many([s, s], [Y, Z], opts)
    finally:
        wa.set_many_shape(0)
    ref = s.solve_batch(Y, opts)
    assert not any(r_cap.converged) and r_cap.iterations == ref.iterations
    assert max(rel(r_cap.demeaned[:, j], ref.demeaned[:, j]) for j in range(Y.shape[1])) <= 1e-13
    assert all(r_zero.converged) and np.all(r_zero.demeaned == 0.0)
