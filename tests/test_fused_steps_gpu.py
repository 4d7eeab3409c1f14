"""The LSMR steps folded into the additive Schwarz apply against the same steps as launches of
their own (crates/schwarz-precond/src/lsmr/bidiag.rs:99-111, 475-485).

A solve folds p~ <- A'u / beta - (theta / rho) p~ into the apply's restriction and the partial
sums of <v,p~>, <v,v>, <p~,p~> into its prolongation.  Each DOF and column goes through the same
operations in the same order either way, and the partial sums cover the same blocks, so the two
forms give the same bits: x, the demeaned output, the iteration counts and the LSMR estimates."""
import numpy as np
import pytest

import within_amd as wa
from within_amd import _lib

pytestmark = pytest.mark.gpu


def set_fused(on):
    return _lib.lib().wx_testing_set_fused_steps(1 if on else 0)


def _design(seed, n, la, lb, lc):
    rng = np.random.default_rng(seed)
    a = np.minimum((rng.pareto(1.1, n) * la / 40).astype(np.int64), la - 1)
    codes = np.asfortranarray(
        np.column_stack([a, rng.integers(0, lb, n), rng.integers(0, lc, n)]).astype(np.uint32))
    return rng, codes


def _same(r1, r2):
    assert list(r1.iterations) == list(r2.iterations)
    assert np.array_equal(r1.x, r2.x)
    assert np.array_equal(r1.demeaned, r2.demeaned)
    assert np.array_equal(np.asarray(r1.residual), np.asarray(r2.residual))


@pytest.fixture(scope="module")
def solver():
    # a 1,200-level middle term: its kept block is exact and dense (upper 64 x 64 blocks on the
    # matrix cores in mixed precision); a small third term and a Zipf dominant term
    rng, codes = _design(61, 300_000, 20_000, 1_200, 50)
    w = rng.uniform(0.5, 2.0, codes.shape[0])
    return wa.Solver(codes, weights=w), rng


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", [1, 5, 16])
def test_folded_steps_give_the_same_bits(gpu, solver, precision, k):
    s, rng = solver
    print(f"[fused] solve_batch precision={precision} k={k}", flush=True)
    Y = np.asfortranarray(np.random.default_rng(100 + k).standard_normal((s.n_obs, k)))
    opts = wa.LsmrOptions(tol=1e-9, precision=precision)
    prev = set_fused(False)
    try:
        sep = s.solve_batch(Y, opts)
    finally:
        set_fused(prev)
    assert prev == 1
    fol = s.solve_batch(Y, opts)
    assert all(fol.converged)
    _same(fol, sep)


def test_folded_steps_give_the_same_bits_in_a_pack(gpu):
    """Packed members (one block-diagonal batch) read their column state through a per-workgroup
    base in both forms."""
    problems = []
    for q in range(4):
        rng, codes = _design(70 + q, 20_000 + 7_000 * q, 1_500, 120, 9)
        w = rng.uniform(0.5, 2.0, codes.shape[0])
        problems.append((wa.Solver(codes, weights=w), np.asfortranarray(rng.standard_normal((codes.shape[0], 3)))))
    solvers, Ys = [s for s, _ in problems], [Y for _, Y in problems]
    opts = wa.LsmrOptions(tol=1e-9, precision="mixed")
    print("[fused] packed solve_many, shape (4, 1, graph, pack)", flush=True)
    wa.set_many_shape(4, 1, True, True)
    try:
        fol = wa.solve_many(solvers, Ys, opts)
        prev = set_fused(False)
        try:
            sep = wa.solve_many(solvers, Ys, opts)
        finally:
            set_fused(prev)
        again = wa.solve_many(solvers, Ys, opts)
    finally:
        wa.set_many_shape(0)
    for a, b, c in zip(fol, sep, again):
        assert all(a.converged)
        _same(a, b)
        _same(a, c)
