"""Solver.reweight: a Solver for the same design under new weights, built from an existing
Solver's device design (copied on the device), must solve exactly as a Solver built from the
codes with those weights -- same preconditioner build, same device data, so the same bits."""
import numpy as np
import pytest
import torch

import within_amd as wa
from within_amd.autograd import project

pytestmark = pytest.mark.gpu


def case(seed, n=40_000):
    rng = np.random.default_rng(seed)
    p = (np.arange(3000) + 1.0) ** -1.1
    p /= p.sum()
    # a Zipf factor, a uniform one and a small one: unsorted terms, the fused sorted term and
    # a permuted observation order all take part
    codes = np.asfortranarray(np.column_stack([rng.choice(3000, n, p=p), rng.integers(0, 400, n),
                                               rng.integers(0, 12, n)]).astype(np.uint32))
    w1 = np.exp(rng.normal(0.0, 0.5, n))
    w2 = np.exp(rng.normal(0.0, 0.7, n))
    Y = np.asfortranarray(rng.standard_normal((n, 5)))
    return codes, w1, w2, Y


def same(a, b):
    assert a.iterations == b.iterations
    assert np.array_equal(a.demeaned, b.demeaned)
    assert np.array_equal(a.x, b.x)


@pytest.mark.parametrize("pc", [None, wa.PreconditionerConfig.Diagonal, wa.PreconditionerConfig.Off])
@pytest.mark.parametrize("base_weighted", [True, False])
def test_reweight_equals_fresh_solver(gpu, pc, base_weighted):
    codes, w1, w2, Y = case(71)
    print(f"[reweight] pc={pc!r} base_weighted={base_weighted}", flush=True)
    opts = wa.LsmrOptions(tol=1e-10)
    base = wa.Solver(codes, weights=w1 if base_weighted else None, preconditioner=pc)
    for w in (w2, None, w1):
        r = base.reweight(w, pc)
        fresh = wa.Solver(codes, weights=w, preconditioner=pc)
        same(r.solve_batch(Y, opts), fresh.solve_batch(Y, opts))
        assert r.n_dofs == fresh.n_dofs and r.n_obs == fresh.n_obs
        np.testing.assert_array_equal(r.solve(Y[:, 1], opts).demeaned, fresh.solve(Y[:, 1], opts).demeaned)
    # the base is untouched by its reweighted copies (and outlives none of them)
    same(base.solve_batch(Y, opts), wa.Solver(codes, weights=w1 if base_weighted else None,
                                              preconditioner=pc).solve_batch(Y, opts))


def test_reweight_with_kept_preconditioner(gpu):
    codes, w1, w2, Y = case(72)
    opts = wa.LsmrOptions(tol=1e-10)
    base = wa.Solver(codes, weights=w1)
    kept = base.preconditioner
    r = base.reweight(w2, kept)
    same(r.solve_batch(Y, opts), wa.Solver(codes, weights=w2, preconditioner=kept).solve_batch(Y, opts))
    # a converged answer for the new weights, whichever preconditioner got it there
    ref = wa.Solver(codes, weights=w2).solve_batch(Y, opts)
    got = r.solve_batch(Y, opts)
    assert all(got.converged)
    for j in range(Y.shape[1]):
        assert np.linalg.norm(got.demeaned[:, j] - ref.demeaned[:, j]) <= 1e-7 * np.linalg.norm(ref.demeaned[:, j])


def test_reweight_many_packs_like_fresh_solvers(gpu):
    """Reweighted Solvers join packed many-problem runs as their own problems."""
    codes, w1, w2, Y = case(73, n=12_000)
    opts = wa.LsmrOptions(tol=1e-9)
    base = wa.Solver(codes, weights=w1)
    solvers = [base.reweight(w2), base.reweight(w1 * 2.0), base]
    wa.set_many_shape(4, 1, True, True)
    try:
        res = wa.solve_many(solvers, [Y, Y, Y], opts)
    finally:
        wa.set_many_shape(0)
    for s, r in zip(solvers, res):
        ref = s.solve_batch(Y, opts)
        assert all(r.converged)
        for j in range(Y.shape[1]):
            assert abs(r.iterations[j] - ref.iterations[j]) <= 1
            assert np.linalg.norm(r.demeaned[:, j] - ref.demeaned[:, j]) <= 1e-6 * np.linalg.norm(ref.demeaned[:, j])


def test_reweight_rejects_slopes_and_bad_weights(gpu):
    rng = np.random.default_rng(74)
    n = 5000
    a = rng.integers(0, 100, n).astype(np.uint32)
    b = rng.integers(0, 20, n).astype(np.uint32)
    s = wa.Solver([wa.Effect(a, True), wa.Effect(b, True, [rng.standard_normal(n)])])
    with pytest.raises(ValueError, match="slope"):
        s.reweight(np.ones(n))
    plain = wa.Solver(np.asfortranarray(np.column_stack([a, b])))
    with pytest.raises(ValueError):
        plain.reweight(np.ones(n - 1))
    bad = np.ones(n)
    bad[7] = -1.0
    with pytest.raises(ValueError):
        plain.reweight(bad)


def test_projection_through_a_reweighted_solver(gpu):
    """project(Z, w, solver) equals project(Z, w, codes) with the same preconditioner config,
    output and both gradients, bit for bit."""
    codes, w1, w2, Y = case(75, n=20_000)
    mask = np.random.default_rng(76).standard_normal(Y.shape)
    base = wa.Solver(codes, weights=w1)
    out = []
    for design in (codes, base):
        Z = torch.tensor(Y, device="cuda", requires_grad=True)
        wt = torch.tensor(w2, device="cuda", requires_grad=True)
        R = project(Z, wt, design, wa.LsmrOptions(tol=1e-10))
        (torch.tensor(mask, device="cuda") * R * R).sum().backward()
        out.append((R.detach().cpu().numpy(), Z.grad.cpu().numpy(), wt.grad.cpu().numpy()))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
