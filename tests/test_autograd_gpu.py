"""GPU tests of the differentiable projection (within_amd.autograd).

Checked against gradients assembled without the library's backward pass -- dense numpy
(lambda = pinv(D'WD) D'g) on small cases, CPU oracle solves at config 5's full size -- and
against central finite differences of the loss along random directions in weight space.
Loss: sum(mask * demeaned^2), so the upstream gradient is g = 2 mask * demeaned."""
import numpy as np
import pytest
import torch

import within_amd as wa
from within_amd.autograd import project
from oracle.oracle import sparse_design
from tests._configs import cfg5_data
from tests._helpers import oracle_schwarz_from_bytes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def small_case(seed, n, la, lb, k):
    rng = np.random.default_rng(seed)
    p = (np.arange(la) + 1.0) ** -1.1
    p /= p.sum()
    codes = np.asfortranarray(np.column_stack([rng.choice(la, n, p=p), rng.integers(0, lb, n)]).astype(np.uint32))
    w = np.exp(rng.normal(0.0, 0.5, n))
    Y = rng.standard_normal((n, k))
    mask = rng.standard_normal((n, k))
    return rng, codes, w, Y, mask


def run(codes, w, Y, mask, tol, need_grad=True):
    Z = torch.tensor(Y, device="cuda", requires_grad=need_grad)
    wt = torch.tensor(w, device="cuda", requires_grad=need_grad)
    R = project(Z, wt, codes, wa.LsmrOptions(tol=tol))
    loss = (torch.tensor(mask, device="cuda") * R * R).sum()
    value = float(loss.detach())
    if need_grad:
        loss.backward()
        return R.detach().cpu().numpy(), value, Z.grad.cpu().numpy(), wt.grad.cpu().numpy()
    return R.detach().cpu().numpy(), value, None, None


def test_projection_gradients_match_dense_adjoint(gpu):
    _, codes, w, Y, mask = small_case(61, 3000, 150, 40, 3)
    R, _, gz, gw = run(codes, w, Y, mask, 1e-12)
    D = sparse_design(codes).toarray()
    G = D.T @ (w[:, None] * D)
    Gp = np.linalg.pinv(G, rcond=1e-12)
    xs = Gp @ (D.T @ (w[:, None] * Y))
    np.testing.assert_allclose(R, Y - D @ xs, atol=1e-9)
    g = 2 * mask * R
    lam = Gp @ (D.T @ g)
    Dl = D @ lam
    gz_ref = g - w[:, None] * Dl
    gw_ref = -(R * Dl).sum(1)
    assert rel(gz, gz_ref) <= 1e-8
    assert rel(gw, gw_ref) <= 1e-8


def fd_check(rng, codes, w, Y, mask):
    """Central differences of the loss along 20 random directions u in w-space (u_i ~ N(0,1),
    relative step h = 1e-6 per coordinate: w -> w * (1 +- h u)) against the directional
    derivative sum_i grad_w_i * w_i * u_i; returns the worst relative mismatch."""
    _, _, _, gw = run(codes, w, Y, mask, 1e-13)
    h = 1e-6
    worst = 0.0
    for _ in range(20):
        u = rng.standard_normal(len(w))
        _, lp, _, _ = run(codes, w * (1 + h * u), Y, mask, 1e-13, need_grad=False)
        _, lm, _, _ = run(codes, w * (1 - h * u), Y, mask, 1e-13, need_grad=False)
        fd = (lp - lm) / (2 * h)
        an = float(np.dot(gw, w * u))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-12))
    return worst


def test_projection_weight_gradient_finite_differences(gpu):
    rng, codes, w, Y, mask = small_case(62, 20000, 1000, 100, 4)
    worst = fd_check(rng, codes, w, Y, mask)
    print(f"worst relative FD mismatch over 20 directions: {worst:.2e}")
    assert worst <= 1e-6


def test_config5_weight_gradient_finite_differences(gpu):
    codes, w, Y, mask = cfg5_data()
    worst = fd_check(np.random.default_rng(55), codes, w, Y, mask)
    print(f"config 5: worst relative FD mismatch over 20 directions: {worst:.2e}")
    assert worst <= 1e-6


def test_config5_gradients_match_cpu_solves(gpu, oracle):
    codes, w, Y, mask = cfg5_data()
    R, _, gz, gw = run(codes, w, Y, mask, 1e-10)
    # gradients assembled from CPU solves with the same preconditioner data:
    # lambda solves D'WD lambda = D'g, i.e. the projection of g / w; D lambda = g/w - its output
    s = wa.Solver(codes, weights=w)
    sch = oracle_schwarz_from_bytes(oracle, s.preconditioner.to_bytes())
    d = oracle.design(codes)
    g = 2 * mask * R
    Yg = np.asfortranarray(g / w[:, None])
    _, Dm, res = oracle.solve_batch(d, Yg, w, prec="schwarz", schwarz=sch, tol=1e-10, threads=8)
    assert all(r.converged for r in res)
    Dl = Yg - Dm
    gz_ref = g - w[:, None] * Dl
    gw_ref = -(R * Dl).sum(1)
    print(f"config 5: grad_z rel {rel(gz, gz_ref):.2e}, grad_w rel {rel(gw, gw_ref):.2e}")
    assert rel(gz, gz_ref) <= 1e-8
    assert rel(gw, gw_ref) <= 1e-8
