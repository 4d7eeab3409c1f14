"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle and closed forms."""
import os
import pickle

import numpy as np
import pytest

import within_amd as wa
from within_amd import _lib
from oracle.oracle import exact_demean, normal_eq_residual, sparse_design
from tests._helpers import oracle_schwarz_from_bytes, parse_precond, whiten
from tests.conftest import cfg1_data

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SMALL = np.load(os.path.join(HERE, "golden", "small_projections.npz"))
CASES = ["two_factor", "three_factor_weighted", "disconnected"]


def small(case):
    w = SMALL[f"{case}__w"] if f"{case}__w" in SMALL.files else None
    return np.asfortranarray(SMALL[f"{case}__codes"]), np.asfortranarray(SMALL[f"{case}__y"]), w


def apply_op(solver, X, adjoint):
    import ctypes as C
    X = np.asfortranarray(X, dtype=np.float64)
    k = X.shape[1]
    rows = solver.n_dofs if adjoint else solver.n_obs
    Y = np.empty((rows, k), order="F")
    _lib.check(_lib.lib().wx_solver_apply_operator(solver._h, 1 if adjoint else 0, X.ctypes.data, k, Y.ctypes.data))
    return Y


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------------------
# operators D, D' (bit-exact known answers, closed forms, adjoint identity)
# ---------------------------------------------------------------------------

def test_operator_known_answers(gpu):
    # crates/within/src/operator/design/tests.rs:6-57
    codes = np.asfortranarray(np.array([[0, 1, 1, 2, 0], [0, 0, 1, 2, 3]], dtype=np.uint32).T)
    s = wa.Solver(codes, preconditioner=wa.PreconditionerConfig.Off)
    Dx = apply_op(s, np.array([[1.0, 2, 3, 10, 20, 30, 40]]).T, False)[:, 0]
    assert Dx.tolist() == [11.0, 12.0, 22.0, 33.0, 41.0]
    Dtr = apply_op(s, np.array([[1.0, 2, 3, 4, 5]]).T, True)[:, 0]
    assert Dtr.tolist() == [6.0, 5.0, 4.0, 3.0, 3.0, 4.0, 5.0]


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("k", [1, 3, 16])
def test_operator_matches_oracle_bitwise(gpu, oracle, weighted, k):
    rng = np.random.default_rng(17 + k)
    n = 5000
    # unsorted dominant factor -> locality sort; Zipf-like head to exercise split levels
    a = np.minimum((rng.pareto(1.1, n) * 3).astype(np.int64), 399)
    codes = np.asfortranarray(np.column_stack([a, rng.integers(0, 37, n), rng.integers(0, 5, n)]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    s = wa.Solver(codes, weights=w, preconditioner=wa.PreconditionerConfig.Off)
    d = oracle.design(codes)
    X = rng.standard_normal((s.n_dofs, k))
    U = rng.standard_normal((n, k))
    sw = None if w is None else np.sqrt(w)
    got = apply_op(s, X, False)
    got_t = apply_op(s, U, True)
    for j in range(k):
        # the gather's per-row sum order is the reference's: bit-exact
        assert np.array_equal(got[:, j], oracle.gather(d, X[:, j], sw)), j
        # the reduction sums each level in code-sorted row order (split levels pairwise)
        np.testing.assert_allclose(got_t[:, j], oracle.scatter(d, U[:, j], sw), rtol=1e-13, atol=1e-12)
    lhs = np.sum(got * U, axis=0)
    rhs = np.sum(X * got_t, axis=0)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-11)


def _pair_design(seed=23, n=200_000):
    """Three factors where the two unsorted ones (B: 600 levels, C: 9) qualify for the fused D'
    (pieces = runs of equal C inside B's work items, at most n / 4 of them)."""
    rng = np.random.default_rng(seed)
    a = np.minimum((rng.pareto(1.1, n) * 30).astype(np.int64), 2999)
    b = rng.integers(0, 600, n)
    c = rng.integers(0, 9, n)
    codes = np.asfortranarray(np.column_stack([a, b, c]).astype(np.uint32))
    # the library's rule, restated: internal rows by A (stable), B's positions in (B, C) order,
    # work items of at most 64 positions per B level, a piece per run of equal C in an item
    rows = np.argsort(a, kind="stable")
    bi, ci = b[rows], c[rows]
    pos = np.lexsort((ci, bi))
    bp, cp = bi[pos], ci[pos]
    starts = np.searchsorted(bp, np.arange(601))
    pieces = 0
    for l in range(600):
        for p0 in range(starts[l], starts[l + 1], 64):
            seg = cp[p0:min(starts[l + 1], p0 + 64)]
            pieces += 1 + int(np.count_nonzero(seg[1:] != seg[:-1]))
    assert pieces <= n // 4
    return rng, codes


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("k", [2, 3, 16])
def test_fused_pair_adjoint_matches_oracle(gpu, oracle, weighted, k):
    """D' of the two unsorted terms in one pass (even widths; odd widths reduce term by term)."""
    rng, codes = _pair_design()
    n = codes.shape[0]
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    s = wa.Solver(codes, weights=w, preconditioner=wa.PreconditionerConfig.Off)
    d = oracle.design(codes)
    sw = None if w is None else np.sqrt(w)
    U = rng.standard_normal((n, k))
    X = rng.standard_normal((s.n_dofs, k))
    got_t = apply_op(s, U, True)
    got = apply_op(s, X, False)
    for j in range(k):
        np.testing.assert_allclose(got_t[:, j], oracle.scatter(d, U[:, j], sw), rtol=1e-13, atol=1e-11)
    np.testing.assert_allclose(np.sum(got * U, axis=0), np.sum(X * got_t, axis=0), rtol=1e-11)
    # bitwise reproducible run to run
    assert np.array_equal(apply_op(s, U, True), got_t)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_pair_solve_matches_exact(gpu, precision):
    rng, codes = _pair_design(seed=29)
    n = codes.shape[0]
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, 4)))
    r = wa.solve_batch(codes, Y, weights=w, options=wa.LsmrOptions(tol=1e-10, precision=precision))
    assert all(r.converged)
    ex = exact_demean(sparse_design(codes), Y, w)
    for j in range(4):
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j


# ---------------------------------------------------------------------------
# full solves
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("prec", ["off", "diagonal"])
def test_solve_matches_oracle_same_algorithm(gpu, oracle, case, prec):
    codes, y, w = small(case)
    pc = wa.PreconditionerConfig.Off if prec == "off" else wa.PreconditionerConfig.Diagonal
    opts = wa.LsmrOptions(tol=1e-10, maxiter=5000)
    r = wa.solve_batch(codes, y, weights=w, options=opts, preconditioner=pc)
    d = oracle.design(codes)
    _, Dm, res = oracle.solve_batch(d, y, w, prec=prec, tol=1e-10, maxiter=5000)
    assert all(r.converged)
    # same algorithm, different FP64 rounding (the device keeps u_1 unnormalised and sums in
    # its own order), amplified by the unpreconditioned iterations: the bar is the north-star
    # tolerance, 1e-8 relative in norm, per column, against the oracle and the exact projection
    ex = SMALL[f"{case}__demeaned"]
    for j in range(y.shape[1]):
        assert rel(r.demeaned[:, j], Dm[:, j]) <= 1e-8, j
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
    np.testing.assert_allclose(r.demeaned, ex, atol=1e-7)
    assert max(abs(a - b.iterations) for a, b in zip(r.iterations, res)) <= 2


@pytest.mark.parametrize("case", CASES)
def test_additive_schwarz_apply_matches_oracle(gpu, oracle, case):
    codes, y, w = small(case)
    s = wa.Solver(codes, weights=w)
    p = s.preconditioner
    assert repr(p).startswith("Preconditioner(Additive")
    sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = rng.standard_normal(s.n_dofs)
        np.testing.assert_allclose(p.apply(x), sch.apply(x), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("case", CASES)
def test_additive_solve_matches_oracle_and_exact(gpu, oracle, case):
    codes, y, w = small(case)
    s = wa.Solver(codes, weights=w)
    r = s.solve_batch(y, wa.LsmrOptions(tol=1e-10))
    sch = oracle_schwarz_from_bytes(oracle, s.preconditioner.to_bytes())
    d = oracle.design(codes)
    # the oracle solves in the internal (locality-sorted) order: permute like the product
    _, Dm, res = oracle.solve_batch(d, y, w, prec="schwarz", schwarz=sch, tol=1e-10)
    assert all(r.converged) and all(q.converged for q in res)
    ex = SMALL[f"{case}__demeaned"]
    for j in range(y.shape[1]):
        assert rel(r.demeaned[:, j], Dm[:, j]) <= 1e-8, j
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
    np.testing.assert_allclose(r.demeaned, Dm, atol=1e-9)
    np.testing.assert_allclose(r.demeaned, ex, atol=1e-8)
    np.testing.assert_allclose(r.true_residual, normal_eq_residual(codes, y, r.demeaned, w), rtol=1e-6, atol=1e-14)


def test_additive_schwarz_is_spd(gpu):
    codes, y, w = small("three_factor_weighted")
    p = wa.Solver(codes, weights=w).preconditioner
    rng = np.random.default_rng(2)
    n = p.nrows
    X = rng.standard_normal((n, 6))
    MX = np.column_stack([p.apply(X[:, j].copy()) for j in range(6)])
    G = X.T @ MX
    np.testing.assert_allclose(G, G.T, rtol=1e-10, atol=1e-12)
    assert np.all(np.linalg.eigvalsh((G + G.T) / 2) > -1e-10)


def test_config1_parity(gpu, oracle):
    """README data, default Additive, tol 1e-8, maxiter 1000."""
    fe, y = cfg1_data()
    s = wa.Solver(fe)
    r = s.solve(y)
    assert r.converged
    # the same algorithm on the CPU (oracle LSMR restatement + this preconditioner's data)
    sch = oracle_schwarz_from_bytes(oracle, s.preconditioner.to_bytes())
    _, Dm, res = oracle.solve_batch(oracle.design(fe), y, prec="schwarz", schwarz=sch, tol=1e-8, maxiter=1000)
    ex = exact_demean(sparse_design(fe), y)
    print(f"[cfg1] iterations gpu={r.iterations} cpu={res[0].iterations} rel(gpu,cpu)={rel(r.demeaned, Dm[:, 0]):.3e} "
          f"rel(gpu,exact)={rel(r.demeaned, ex):.3e} true_residual={r.true_residual:.3e}", flush=True)
    assert rel(r.demeaned, Dm[:, 0]) <= 1e-8
    assert rel(r.demeaned, ex) <= 1e-7
    np.testing.assert_allclose(r.true_residual, normal_eq_residual(fe, y, r.demeaned)[0], rtol=1e-6, atol=1e-13)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_columnwise_and_is_deterministic(gpu, precision):
    rng = np.random.default_rng(9)
    n = 20000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 900, n), rng.integers(0, 60, n),
                                               rng.integers(0, 8, n)]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, 5)))
    s = wa.Solver(codes, weights=w)
    opts = wa.LsmrOptions(tol=1e-10, precision=precision)
    b1 = s.solve_batch(Y, opts)
    b2 = s.solve_batch(Y, opts)
    assert np.array_equal(b1.demeaned, b2.demeaned) and np.array_equal(b1.x, b2.x)
    for j in [0, 3]:
        r = s.solve(Y[:, j].copy(), opts)
        # a single column reproduces its batch column (same per-column arithmetic)
        np.testing.assert_allclose(r.demeaned, b1.demeaned[:, j], rtol=0, atol=1e-9)
    ex = exact_demean(sparse_design(codes), Y, w)
    assert rel(b1.demeaned, ex) <= 1e-8


def test_edge_cases(gpu):
    rng = np.random.default_rng(4)
    n = 3000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 50, n), rng.integers(0, 7, n)]).astype(np.uint32))
    s = wa.Solver(codes)
    # zero RHS -> ZeroRhs, x = 0
    r = s.solve(np.zeros(n))
    assert r.converged and r.iterations == 0 and r.stop_reason == "ZeroRhs" and not r.x.any()
    # y in the column space of D -> demeaned ~ 0
    xt = rng.standard_normal(s.n_dofs)
    D = sparse_design(codes)
    r = s.solve(D @ xt, wa.LsmrOptions(tol=1e-12))
    assert np.abs(r.demeaned).max() < 1e-8
    # maxiter exhausted (unpreconditioned: the exact two-factor Schwarz converges in one step)
    s_off = wa.Solver(codes, preconditioner=wa.PreconditionerConfig.Off)
    r = s_off.solve(rng.standard_normal(n), wa.LsmrOptions(tol=1e-14, maxiter=2))
    assert not r.converged and r.iterations == 2 and r.stop_reason == "MaxIterations"
    # empty batch
    b = s.solve_batch(np.zeros((n, 0)))
    assert b.x.shape == (s.n_dofs, 0) and b.demeaned.shape == (n, 0)
    # non-finite response -> ValueError
    yb = rng.standard_normal(n)
    yb[7] = np.nan
    with pytest.raises(ValueError, match="finite"):
        s.solve(yb)
    # zero weights exclude rows (edge_cases.rs:74-160): equal to dropping them
    w = np.ones(n)
    w[:500] = 0.0
    y = rng.standard_normal(n)
    r = wa.solve(codes, y, weights=w, options=wa.LsmrOptions(tol=1e-11))
    ex = exact_demean(D, y, w)
    np.testing.assert_allclose(r.demeaned[500:], ex[500:], atol=1e-8)
    # uniform weights == unweighted (edge_cases.rs:323-351)
    r1 = wa.solve(codes, y, weights=np.full(n, 2.5), options=wa.LsmrOptions(tol=1e-11))
    r0 = wa.solve(codes, y, options=wa.LsmrOptions(tol=1e-11))
    np.testing.assert_allclose(r1.demeaned, r0.demeaned, atol=1e-8)


def test_locality_sort_is_transparent(gpu):
    rng = np.random.default_rng(12)
    n = 8000
    codes = np.column_stack([rng.integers(0, 400, n), rng.integers(0, 30, n)]).astype(np.uint32)
    y = rng.standard_normal(n)
    order = np.argsort(codes[:, 0], kind="stable")
    r_uns = wa.solve(np.asfortranarray(codes), y, options=wa.LsmrOptions(tol=1e-11))
    r_srt = wa.solve(np.asfortranarray(codes[order]), y[order], options=wa.LsmrOptions(tol=1e-11))
    np.testing.assert_allclose(r_uns.demeaned[order], r_srt.demeaned, atol=1e-8)
    with pytest.warns(UserWarning, match="F-contiguous"):
        wa.solve(np.ascontiguousarray(codes), y)


def test_preconditioner_pickle_round_trip(gpu):
    codes, y, w = small("three_factor_weighted")
    s = wa.Solver(codes, weights=w)
    p = s.preconditioner
    q = pickle.loads(pickle.dumps(p))
    assert repr(q) == repr(p) and q.nrows == p.nrows
    x = np.random.default_rng(0).standard_normal(p.nrows)
    assert np.array_equal(p.apply(x), q.apply(x))
    s2 = wa.Solver(codes, weights=w, preconditioner=q)
    a = s.solve_batch(y)
    b = s2.solve_batch(y)
    assert np.array_equal(a.demeaned, b.demeaned)
    with pytest.raises(ValueError, match="dimension"):
        wa.Solver(codes[:, :2].copy(order="F"), preconditioner=q)
    d = wa.Solver(codes, preconditioner=wa.PreconditionerConfig.Diagonal).preconditioner
    assert repr(d).startswith("Preconditioner(Diagonal") and parse_precond(d.to_bytes())["kind"] == 0
    assert wa.Solver(codes, preconditioner=wa.PreconditionerConfig.Off).preconditioner is None


def test_single_factor_runs_plain_lsmr(gpu):
    # one factor: no factor pair, the Additive build falls back to plain LSMR (schwarz.rs:249-252)
    rng = np.random.default_rng(8)
    codes = np.asfortranarray(rng.integers(0, 40, (2000, 1)).astype(np.uint32))
    y = rng.standard_normal(2000)
    s = wa.Solver(codes)
    assert s.preconditioner is None
    r = s.solve(y, wa.LsmrOptions(tol=1e-12))
    means = np.bincount(codes[:, 0], weights=y) / np.bincount(codes[:, 0])
    np.testing.assert_allclose(r.demeaned, y - means[codes[:, 0]], atol=1e-9)


@pytest.mark.parametrize("case", CASES)
def test_sparse_factor_apply_matches_oracle(gpu, oracle, case):
    # dense_threshold=0 routes every kept block through approximate Cholesky + level-scheduled solves
    codes, y, w = small(case)
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    s = wa.Solver(codes, weights=w, preconditioner=cfg)
    p = s.preconditioner
    sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
    assert (parse_precond(p.to_bytes())["fac_kind"] == 1).all()
    rng = np.random.default_rng(3)
    for _ in range(3):
        x = rng.standard_normal(s.n_dofs)
        np.testing.assert_allclose(p.apply(x), sch.apply(x), rtol=1e-11, atol=1e-12)
    for precision in ["fp64", "mixed"]:
        r = s.solve_batch(y, wa.LsmrOptions(tol=1e-10, precision=precision))
        assert all(r.converged)
        np.testing.assert_allclose(r.demeaned, SMALL[f"{case}__demeaned"], atol=1e-8)


@pytest.mark.parametrize("rows", ["uniform", "zipf"])
def test_large_dense_kept_block_is_the_schur_pseudo_inverse(gpu, rows):
    """Kept blocks above 1,024 levels get their Schur complement S = D_k - C' D_e^-1 C assembled
    and pseudo-inverted on the GPU (rocSOLVER Cholesky of the grounded minor, projected).  The
    stored S+ must be numpy's pseudo-inverse of S rebuilt from the exported C and D_e^-1.  With
    Zipf-distributed eliminated levels the largest ones touch most kept levels, and their terms
    come from the dense product on the matrix cores instead of the row kernel."""
    import scipy.sparse as sp
    rng = np.random.default_rng(33)
    n = 60_000
    if rows == "uniform":
        a = rng.integers(0, 3000, n)
    else:
        pz = (np.arange(3000) + 1.0) ** -1.1
        a = rng.choice(3000, n, p=pz / pz.sum())
    codes = np.asfortranarray(np.column_stack([a, rng.integers(0, 1500, n)]).astype(np.uint32))
    s = wa.Solver(codes, weights=rng.uniform(0.5, 2.0, n))
    p = parse_precond(s.preconditioner.to_bytes())
    assert list(p["fac_kind"]) == [0] and int(p["nk"][0]) == 1500 and p["gauge"][0] == 0
    ne, nk = int(p["ne"][0]), int(p["nk"][0])
    cptr = p["c_ptr"][: ne + 1].astype(np.int64)
    C = sp.csr_matrix((p["c_val"], p["c_idx"].astype(np.int64), cptr), shape=(ne, nk))
    S = -(C.T @ sp.diags(p["inv_de"][:ne]) @ C).toarray()
    np.fill_diagonal(S, 0.0)
    np.fill_diagonal(S, -S.sum(axis=1))  # a Laplacian: rows sum to zero
    Sp = p["dense"][int(p["dense_off"][0]):int(p["dense_off"][0]) + nk * nk].reshape(nk, nk)
    ref = np.linalg.pinv(S, rcond=1e-12, hermitian=True)
    np.testing.assert_allclose(Sp, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())
    np.testing.assert_allclose(S @ Sp @ S, S, rtol=1e-7, atol=1e-9 * np.abs(S).max())


def test_sparse_factor_dense_tail_matches_oracle(gpu, oracle, capfd, monkeypatch):
    """A worker-firm panel (Zipf firms) eliminates the hub firms last, in a chain of one-row
    levels; the device solves that tail as one dense product M_T = (I+L_TT')^-1 D_T^+ (I+L_TT)^-1.
    The apply must match the oracle's level-by-level restatement of the same factor."""
    monkeypatch.setenv("WITHIN_AMD_VERBOSE", "1")
    rng = np.random.default_rng(31)
    nw, T, nf = 20_000, 5, 3_000
    p = (np.arange(nf) + 1.0) ** -1.2
    p /= p.sum()
    firm = np.empty((nw, T), np.int64)
    firm[:, 0] = rng.choice(nf, nw, p=p)
    for t in range(1, T):
        move = rng.uniform(size=nw) < 0.05
        firm[:, t] = np.where(move, rng.integers(0, nf, nw), firm[:, t - 1])
    codes = np.asfortranarray(np.column_stack([np.repeat(np.arange(nw), T), firm.ravel(),
                                               np.tile(np.arange(T), nw)]).astype(np.uint32))
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    s = wa.Solver(codes, preconditioner=cfg)
    pc = s.preconditioner
    x = rng.standard_normal(s.n_dofs)
    got = pc.apply(x)
    assert "dense tail" in capfd.readouterr().err, "the factor has no dense tail: the test does not cover it"
    sch = oracle_schwarz_from_bytes(oracle, pc.to_bytes())
    np.testing.assert_allclose(got, sch.apply(x), rtol=1e-10, atol=1e-11)
    y = rng.standard_normal((codes.shape[0], 2))
    ex = exact_demean(sparse_design(codes), y)
    for precision in ["fp64", "mixed"]:
        r = s.solve_batch(np.asfortranarray(y), wa.LsmrOptions(tol=1e-10, precision=precision))
        assert all(r.converged)
        assert rel(r.demeaned, ex) <= 1e-8


# ---------------------------------------------------------------------------
# additive Schwarz over slope (signed) channel pairs
# ---------------------------------------------------------------------------

def _slope_design(seed, n=3000, la=80, lb=30, lc=4):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, la, n).astype(np.uint32)
    b = rng.integers(0, lb, n).astype(np.uint32)
    c = rng.integers(0, lc, n).astype(np.uint32)
    z = rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    design = [wa.Effect(a, True), wa.Effect(b, True, [z]), wa.Effect(c, True)]
    D = sparse_design(np.column_stack([a, b, c]), effects=[(True, []), (True, [z]), (True, [])])
    # the solver's columns: the slope whitened per level (its normal-equation residual basis)
    Dw = sparse_design(np.column_stack([a, b, c]), effects=[(True, []), (True, [whiten(b, z, w)]), (True, [])])
    return rng, design, D, Dw, w


def _local(dense_threshold):
    """None: the default local solver (device dense policy); 0: approximate Cholesky for every
    kept block (the reference rule with threshold 0)."""
    if dense_threshold is None:
        return wa.config.LocalSolverConfig()
    return wa.config.LocalSolverConfig(dense_threshold=dense_threshold, dense_policy="reference")


@pytest.mark.parametrize("dense_threshold", [None, 0])
def test_slope_schwarz_apply_matches_oracle(gpu, oracle, dense_threshold):
    rng, design, _, _, w = _slope_design(51)
    cfg = wa.config.AdditiveSchwarz(_local(dense_threshold))
    s = wa.Solver(design, weights=w, preconditioner=cfg)
    p = s.preconditioner
    kinds = set(parse_precond(p.to_bytes())["fac_kind"].tolist())
    assert kinds == ({0} if dense_threshold is None else {1, 2})
    sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
    for _ in range(3):
        x = rng.standard_normal(s.n_dofs)
        np.testing.assert_allclose(p.apply(x), sch.apply(x), rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("dense_threshold", [None, 0])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_slope_additive_solve_matches_exact(gpu, dense_threshold, precision):
    rng, design, D, Dw, w = _slope_design(52)
    Y = np.asfortranarray(rng.standard_normal((D.shape[0], 3)))
    cfg = wa.config.AdditiveSchwarz(_local(dense_threshold))
    s = wa.Solver(design, weights=w, preconditioner=cfg)
    r = s.solve_batch(Y, wa.LsmrOptions(tol=1e-10, precision=precision))
    assert all(r.converged)
    ex = exact_demean(D, Y, w)
    for j in range(Y.shape[1]):
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
    np.testing.assert_allclose(r.true_residual, normal_eq_residual(Dw, Y, r.demeaned, w), rtol=1e-6, atol=1e-13)
    # the preconditioner does its job: far fewer iterations than the unpreconditioned solve
    off = wa.Solver(design, weights=w, preconditioner=wa.PreconditionerConfig.Off)
    r0 = off.solve_batch(Y, wa.LsmrOptions(tol=1e-10, maxiter=20000))
    assert max(r.iterations) * 3 < max(r0.iterations)


def test_worker_firm_panel_with_firm_slope(gpu):
    # a miniature of the bipartite worker-firm panel: workers x 5 periods, Zipf initial firm,
    # moves w.p. 0.05, year effect, firm intercept + slope; tol 1e-10 against the exact projection
    rng = np.random.default_rng(3)
    nw, T, nf = 3000, 5, 400
    p = (np.arange(nf) + 1.0) ** -1.2
    p /= p.sum()
    firm = np.empty((nw, T), np.int64)
    firm[:, 0] = rng.choice(nf, nw, p=p)
    for t in range(1, T):
        move = rng.uniform(size=nw) < 0.05
        firm[:, t] = np.where(move, rng.integers(0, nf, nw), firm[:, t - 1])
    worker = np.repeat(np.arange(nw), T).astype(np.uint32)
    firm = firm.ravel().astype(np.uint32)
    year = np.tile(np.arange(T), nw).astype(np.uint32)
    z = rng.standard_normal(nw * T)
    design = [wa.Effect(worker, True), wa.Effect(firm, True, [z]), wa.Effect(year, True)]
    D = sparse_design(np.column_stack([worker, firm, year]), effects=[(True, []), (True, [z]), (True, [])])
    Y = np.asfortranarray(rng.standard_normal((nw * T, 4)))
    r = wa.solve_batch(design, Y, options=wa.LsmrOptions(tol=1e-10))
    assert all(r.converged)
    ex = exact_demean(D, Y)
    for j in range(4):
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j


# ---------------------------------------------------------------------------
# windowed reorthogonalisation (local_size) on the device
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("k", [8, 16, 64])
def test_wide_batch_with_small_factor_matches_oracle(gpu, oracle, k):
    """Wide even batches run the wavefront gather (4, 8 and 32 lanes per row: 16, 8 and 2 rows
    per wave step) over a Zipf-sorted term with split and merged levels."""
    rng = np.random.default_rng(80 + k)
    n = 60_000
    p = (np.arange(3000) + 1.0) ** -1.1
    p /= p.sum()
    codes = np.asfortranarray(np.column_stack([rng.choice(3000, n, p=p), rng.integers(0, 400, n),
                                               rng.integers(0, 12, n)]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, k)))
    print(f"[wide batch] k={k}", flush=True)
    r = wa.solve_batch(codes, Y, weights=w, options=wa.LsmrOptions(tol=1e-10, maxiter=5000),
                       preconditioner=wa.PreconditionerConfig.Diagonal)
    _, Dm, res = oracle.solve_batch(oracle.design(codes), Y, w, prec="diagonal", tol=1e-10, maxiter=5000, threads=8)
    assert all(r.converged) and all(q.converged for q in res)
    ex = exact_demean(sparse_design(codes), Y, w)
    for j in range(k):
        assert rel(r.demeaned[:, j], Dm[:, j]) <= 1e-8, j
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
    assert max(abs(a - b.iterations) for a, b in zip(r.iterations, res)) <= 2


ZW_LIVE = [2] + list(range(5, 16))


def test_wide_batch_with_stopped_leading_columns(gpu, oracle):
    """Columns 0 and 1 are zero (stopped before the first iteration), so the wavefront gather's
    lanes of that column pair never enter its loop, and the other pairs' level sums must not
    depend on them; column 4 is zero beside a live column 5, and column 3 (which the fixed
    effects fit exactly) runs longer than most."""
    rng = np.random.default_rng(97)
    n, k = 60_000, 16
    p = (np.arange(3000) + 1.0) ** -1.1
    p /= p.sum()
    codes = np.asfortranarray(np.column_stack([rng.choice(3000, n, p=p), rng.integers(0, 400, n),
                                               rng.integers(0, 12, n)]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, k)))
    Y[:, 0] = Y[:, 1] = Y[:, 4] = 0.0
    Y[:, 3] = rng.standard_normal(3000)[codes[:, 0]] + rng.standard_normal(400)[codes[:, 1]]
    print(f"[wide batch, stopped leading columns] k={k}", flush=True)
    opts = wa.LsmrOptions(tol=1e-10, maxiter=5000)
    for pc in (wa.PreconditionerConfig.Diagonal, None):
        r = wa.solve_batch(codes, Y, weights=w, options=opts, preconditioner=pc)
        assert all(r.converged)
        assert r.iterations[0] == r.iterations[1] == r.iterations[4] == 0
        ex = exact_demean(sparse_design(codes), Y, w)
        for j in (0, 1, 4):
            assert np.array_equal(r.demeaned[:, j], np.zeros(n))
        assert np.abs(r.demeaned[:, 3]).max() <= 1e-7 * np.abs(Y[:, 3]).max()
        for j in ZW_LIVE:
            assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
        # each column as its own solve
        for j in (2, 7, 15):
            one = wa.solve_batch(codes, Y[:, [j]], weights=w, options=opts, preconditioner=pc)
            assert rel(r.demeaned[:, j], one.demeaned[:, 0]) <= 1e-9, j
    _, Dm, res = oracle.solve_batch(oracle.design(codes), Y, w, prec="diagonal", tol=1e-10, maxiter=5000, threads=8)
    r = wa.solve_batch(codes, Y, weights=w, options=opts, preconditioner=wa.PreconditionerConfig.Diagonal)
    for j in ZW_LIVE:
        assert rel(r.demeaned[:, j], Dm[:, j]) <= 1e-8, j


@pytest.mark.parametrize("prec", ["off", "diagonal", "schwarz"])
@pytest.mark.parametrize("local_size", [1, 5, 40])
def test_local_size_matches_oracle(gpu, oracle, prec, local_size):
    codes, y, w = small("three_factor_weighted")
    opts = wa.LsmrOptions(tol=1e-10, maxiter=5000, local_size=local_size)
    d = oracle.design(codes)
    if prec == "schwarz":
        s = wa.Solver(codes, weights=w)
        r = s.solve_batch(y, opts)
        sch = oracle_schwarz_from_bytes(oracle, s.preconditioner.to_bytes())
        _, Dm, res = oracle.solve_batch(d, y, w, prec="schwarz", schwarz=sch, tol=1e-10, local_size=local_size)
    else:
        pc = wa.PreconditionerConfig.Off if prec == "off" else wa.PreconditionerConfig.Diagonal
        r = wa.solve_batch(codes, y, weights=w, options=opts, preconditioner=pc)
        _, Dm, res = oracle.solve_batch(d, y, w, prec=prec, tol=1e-10, maxiter=5000, local_size=local_size)
    assert all(r.converged) and all(q.converged for q in res)
    ex = SMALL["three_factor_weighted__demeaned"]
    for j in range(y.shape[1]):
        assert rel(r.demeaned[:, j], Dm[:, j]) <= 1e-8, j
        assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8, j
    assert max(abs(a - b.iterations) for a, b in zip(r.iterations, res)) <= 2


def test_local_size_mixed_precision_and_batch(gpu):
    rng = np.random.default_rng(71)
    n = 6000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 300, n), rng.integers(0, 40, n),
                                               rng.integers(0, 6, n)]).astype(np.uint32))
    Y = np.asfortranarray(rng.standard_normal((n, 5)))
    s = wa.Solver(codes)
    ex = exact_demean(sparse_design(codes), Y)
    for precision in ("fp64", "mixed"):
        r = s.solve_batch(Y, wa.LsmrOptions(tol=1e-10, local_size=8, precision=precision))
        assert all(r.converged)
        for j in range(Y.shape[1]):
            assert rel(r.demeaned[:, j], ex[:, j]) <= 1e-8
    # batched columns equal column-by-column solves with the same window
    rb = s.solve_batch(Y, wa.LsmrOptions(tol=1e-10, local_size=8))
    for j in range(2):
        r1 = s.solve_batch(np.asfortranarray(Y[:, j:j + 1]), wa.LsmrOptions(tol=1e-10, local_size=8))
        np.testing.assert_allclose(r1.demeaned[:, 0], rb.demeaned[:, j], rtol=0, atol=1e-12)


def _apply_prec(p, X, precision):
    """M^{-1} X (n x k) with the arithmetic of a solve of `precision` (0 FP64, 1 mixed)."""
    X = np.asfortranarray(X, dtype=np.float64)
    Y = np.empty_like(X)
    _lib.check(_lib.lib().wx_testing_precond_apply(p._h, X.ctypes.data, X.shape[1], Y.ctypes.data, precision))
    return Y


@pytest.mark.parametrize("kind", ["dense", "dense_tiled", "sparse", "slope_cover"])
def test_mixed_precision_apply_is_symmetric_and_definite(gpu, kind):
    """The preconditioner of a mixed-precision solve (FP32 coefficients and FP32 slot-space
    vectors, FP64 accumulation) is applied as an explicit matrix: it stays symmetric and
    positive semi-definite to FP32 rounding -- far inside the sqrt(eps) margin of the
    indefiniteness check (crates/schwarz-precond/src/lsmr/bidiag.rs:99-111) -- and within FP32
    rounding of the FP64 apply."""
    rng = np.random.default_rng(71)
    n = 6000
    if kind == "dense_tiled":
        # a 300-level kept block: the FP32 run applies its upper triangle of 64 x 64 blocks on
        # the matrix cores (five block rows, the last one 44 rows, two super-rows of four)
        n = 20000
        a, b, c = rng.integers(0, 600, n), rng.integers(0, 300, n), rng.integers(0, 6, n)
        design = np.asfortranarray(np.column_stack([a, b, c]).astype(np.uint32))
        w = rng.uniform(0.5, 2.0, n)
        cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=1024))
    elif kind == "slope_cover":
        _, design, _, _, w = _slope_design(53)
        cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    else:
        a, b, c = rng.integers(0, 300, n), rng.integers(0, 40, n), rng.integers(0, 6, n)
        design = np.asfortranarray(np.column_stack([a, b, c]).astype(np.uint32))
        w = rng.uniform(0.5, 2.0, n)
        ls = (wa.config.LocalSolverConfig() if kind == "dense"
              else wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
        cfg = wa.config.AdditiveSchwarz(ls)
    s = wa.Solver(design, weights=w, preconditioner=cfg)
    p = s.preconditioner
    N = s.n_dofs
    E = np.eye(N)
    M32 = np.column_stack([_apply_prec(p, E[:, c0:c0 + 16], 1) for c0 in range(0, N, 16)])
    M64 = np.column_stack([_apply_prec(p, E[:, c0:c0 + 16], 0) for c0 in range(0, N, 16)])
    np.testing.assert_allclose(M64[:, 0], p.apply(np.ascontiguousarray(E[:, 0])), rtol=0,
                               atol=1e-14 * np.abs(M64).max())
    scale = np.abs(M64).max()
    assert np.abs(M32 - M64).max() <= 1e-5 * scale
    assert np.abs(M32 - M32.T).max() <= 1e-5 * scale
    ev = np.linalg.eigvalsh(0.5 * (M32 + M32.T))
    assert ev.min() >= -1e-5 * ev.max()
    # more than 16 columns: the products run in 16-column windows
    X = rng.standard_normal((N, 21))
    np.testing.assert_allclose(_apply_prec(p, X, 1), M64 @ X, rtol=0, atol=1e-5 * np.abs(M64 @ X).max())
    # the FP32 vectors change only rounding: a mixed solve reaches the FP64 solve's answer
    Y = np.asfortranarray(rng.standard_normal((s.n_obs, 2)))
    rm = s.solve_batch(Y, wa.LsmrOptions(tol=1e-10, precision="mixed"))
    rf = s.solve_batch(Y, wa.LsmrOptions(tol=1e-10))
    assert all(rm.converged) and all(rf.converged)
    assert max(rel(rm.demeaned[:, j], rf.demeaned[:, j]) for j in range(2)) <= 1e-8


def test_dense_tails_stay_bounded_on_many_mid_size_components(gpu, oracle):
    """Approximate Cholesky everywhere (dense threshold 0) on a design of 40 disconnected blocks:
    many mid-size sparse kept blocks.  The apply matches the oracle's level-by-level restatement of
    the same factors, and the dense tails' entries stay within 8 times the factors' entries and
    rows (the device forms a tail only where it does not outweigh the factor it replaces)."""
    import re
    import subprocess
    import sys
    code = r'''
import numpy as np, within_amd as wa
rng = np.random.default_rng(91)
blocks = []
for b in range(40):
    n = 3000
    a = rng.integers(0, 150, n) + 150 * b
    c = rng.integers(0, 70, n) + 70 * b
    blocks.append(np.column_stack([a, c]))
codes = np.asfortranarray(np.concatenate(blocks).astype(np.uint32))
np.save("gpurun_out/_tails_codes.npy", codes)
cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
s = wa.Solver(codes, preconditioner=cfg)
p = s.preconditioner
open("gpurun_out/_tails_pre.bin", "wb").write(p.to_bytes())
x = np.random.default_rng(5).standard_normal(s.n_dofs)
np.save("gpurun_out/_tails_x.npy", x)
np.save("gpurun_out/_tails_y.npy", p.apply(x))
r = s.solve(np.random.default_rng(6).standard_normal(codes.shape[0]), wa.LsmrOptions(tol=1e-10))
print("converged", r.converged, r.iterations)
'''
    import os
    os.makedirs("gpurun_out", exist_ok=True)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                         env={**os.environ, "WITHIN_AMD_VERBOSE": "1"})
    assert out.returncode == 0, out.stderr[-3000:]
    assert "converged True" in out.stdout
    m = re.findall(r"sparse factors: (\d+) rows, (\d+) factor entries, dense tails of (\d+) rows \((\d+) entries",
                   out.stderr)
    assert m, out.stderr[-2000:]
    rows, entries, t_rows, t_entries = map(int, m[-1])
    print(f"factors: {rows} rows, {entries} entries; tails: {t_rows} rows, {t_entries} entries")
    assert t_entries <= 8 * (entries + rows)
    sch = oracle_schwarz_from_bytes(oracle, open("gpurun_out/_tails_pre.bin", "rb").read())
    x = np.load("gpurun_out/_tails_x.npy")
    assert rel(np.load("gpurun_out/_tails_y.npy"), sch.apply(x)) <= 1e-11


@pytest.mark.parametrize("case", ["weighted_3f", "unweighted_2f_gaps", "zero_weights", "large_disconnected"])
def test_device_cross_tabs_give_the_host_preconditioner(gpu, case):
    """The plain pairs' cross-tabs built on the device (stable sort by cell, each cell's weights
    added in observation order) make the preconditioner byte for byte equal to the host-built one
    (WITHIN_AMD_HOST_SETUP=1), for the default local solver and for approximate Cholesky
    everywhere."""
    from within_amd._within import build_preconditioner
    rng = np.random.default_rng(61)
    n = 40_000
    if case == "weighted_3f":
        codes = np.column_stack([rng.integers(0, 2000, n), rng.integers(0, 300, n), rng.integers(0, 9, n)])
        w = rng.uniform(0.5, 2.0, n)
    elif case == "unweighted_2f_gaps":
        codes = np.column_stack([rng.choice(np.r_[0:3000:3], n), rng.integers(0, 120, n)])
        w = None
    elif case == "zero_weights":
        codes = np.column_stack([rng.integers(0, 1500, n), rng.integers(0, 200, n), rng.integers(0, 4, n)])
        w = rng.uniform(0.5, 2.0, n)
        w[rng.uniform(size=n) < 0.1] = 0.0
    else:
        # 1.5M rows in two disconnected blocks of levels plus isolated zero-weight levels: enough
        # cells for the parallel transposes and union-find, several components per pair
        n = 1_500_000
        half = rng.uniform(size=n) < 0.6
        a = np.where(half, rng.integers(0, 100_000, n), rng.integers(100_000, 150_000, n))
        b = np.where(half, rng.integers(0, 2_500, n), rng.integers(2_500, 4_000, n))
        c = np.where(half, rng.integers(0, 6, n), rng.integers(6, 9, n))
        codes = np.column_stack([a, b, c])
        w = rng.uniform(0.5, 2.0, n)
        w[a % 997 == 5] = 0.0
    codes = np.asfortranarray(codes.astype(np.uint32))
    for ls in (wa.config.LocalSolverConfig(), wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference")):
        cfg = wa.config.AdditiveSchwarz(ls)
        dev = build_preconditioner(codes, w, cfg).to_bytes()
        os.environ["WITHIN_AMD_HOST_SETUP"] = "1"
        try:
            host = build_preconditioner(codes, w, cfg).to_bytes()
        finally:
            del os.environ["WITHIN_AMD_HOST_SETUP"]
        pd, ph = parse_precond(dev), parse_precond(host)
        if not np.any((pd["fac_kind"] == 0) & (pd["nk"] > 1024)):
            assert dev == host
        else:
            # kept blocks above 1,024 levels are pseudo-inverted on the GPU by the device build
            # and on the host by the host build: equal to rounding; everything else bit for bit
            for f in pd:
                if f != "dense":
                    assert np.array_equal(pd[f], ph[f]), f
            scale = np.abs(ph["dense"]).max()
            assert np.max(np.abs(pd["dense"] - ph["dense"])) <= 1e-9 * scale
    # and the Solver (which builds through the same path) solves with it
    r = wa.solve(codes, np.asfortranarray(rng.standard_normal(n)), weights=w)
    assert r.converged


def test_true_residual_can_be_skipped(gpu):
    """LsmrOptions(true_residual=False) leaves the reference's work only: the same x, demeaned
    output and LSMR estimate bit for bit, with the FP64 true residual reported as NaN."""
    rng = np.random.default_rng(71)
    n = 30_000
    a0 = rng.integers(0, 900, n)
    designs = [
        np.column_stack([a0, rng.integers(0, 40, n)]),                          # rows permuted
        np.column_stack([np.sort(a0), rng.integers(0, 40, n), rng.integers(0, 7, n)]),  # caller order kept
    ]
    w = rng.uniform(0.5, 2.0, n)
    for codes in designs:
        s = wa.Solver(np.asfortranarray(codes.astype(np.uint32)), weights=w)
        for k in (4, 3):
            Y = np.asfortranarray(rng.standard_normal((n, k)))
            for precision in ("fp64", "mixed"):
                a = s.solve_batch(Y, wa.LsmrOptions(precision=precision))
                b = s.solve_batch(Y, wa.LsmrOptions(precision=precision, true_residual=False))
                assert np.array_equal(a.x, b.x) and np.array_equal(a.demeaned, b.demeaned)
                assert list(a.residual) == list(b.residual) and list(a.iterations) == list(b.iterations)
                assert all(np.isfinite(a.true_residual)) and all(np.isnan(b.true_residual))


def test_device_built_large_dense_block_applies_like_the_host_built_one(gpu, oracle):
    """A 1,500-level kept block: its Schur complement and pseudo-inverse built on the GPU (row
    kernel, the heavy eliminated levels' DGEMM, rocSOLVER) against the host build of the same
    design (WITHIN_AMD_HOST_SETUP=1: host cross-tabs, host Schur complement and Cholesky).  The
    two preconditioners share every subdomain, weight and cross block bit for bit, and their
    applies, both by the oracle from the exported bytes, agree to 1e-11."""
    from within_amd._within import build_preconditioner
    rng = np.random.default_rng(91)
    n = 60_000
    pz = (np.arange(3000) + 1.0) ** -1.1
    codes = np.asfortranarray(np.column_stack([rng.choice(3000, n, p=pz / pz.sum()),
                                               rng.integers(0, 1500, n)]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    dev = build_preconditioner(codes, w).to_bytes()
    os.environ["WITHIN_AMD_HOST_SETUP"] = "1"
    try:
        host = build_preconditioner(codes, w).to_bytes()
    finally:
        del os.environ["WITHIN_AMD_HOST_SETUP"]
    pd, ph = parse_precond(dev), parse_precond(host)
    assert int(pd["nk"].max()) == 1500
    for f in ("slot_start", "ne", "nk", "fac_kind", "gauge", "glob", "pw", "fold", "inv_de", "c_ptr", "c_idx", "c_val"):
        assert np.array_equal(pd[f], ph[f]), f
    sd, sh = oracle_schwarz_from_bytes(oracle, dev), oracle_schwarz_from_bytes(oracle, host)
    x = np.random.default_rng(3).standard_normal(int(pd["n"]))
    yd, yh = sd.apply(x), sh.apply(x)
    assert rel(yd, yh) <= 1e-11, rel(yd, yh)
