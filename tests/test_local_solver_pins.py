"""The reference's unit-test pins of its local solver, Schur sampler and approximate Cholesky,
restated against this library's host builder (CPU only).

The reference builds single bipartite blocks by hand (`SddmMatrix::from_dense_for_test`,
`LocalComponent::general_for_test`); here `wx_testing_block_precond` and
`wx_testing_reduced_laplacian` (include/within_amd_testing.h) run the builder's own orientation,
SDDM conversion, Schur reduction and factorisation on such a block.  The local solve is applied
by the oracle's restatement of the device apply, fed the library's exported factors (that apply
is pinned against the device separately)."""
import ctypes as C

import numpy as np
import pytest

import within_amd as wa
from tests._helpers import oracle_schwarz_from_bytes, parse_precond
from within_amd import _lib
from within_amd._within import Preconditioner, _native_config

SPARSE_EXACT = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(
    schur=wa.config.Schur.exact(), dense_threshold=0, dense_policy="reference"))


def _cfg(seed=0, split=1, exact=False, dense_threshold=0, chol_seed=0):
    ls = wa.config.LocalSolverConfig(
        approx_chol=wa.config.ApproxCholConfig(seed=chol_seed, split_merge=2),
        schur=wa.config.Schur.exact() if exact else wa.config.Schur.approximate(
            wa.config.ApproxSchurConfig(seed=seed, split=split)),
        dense_threshold=dense_threshold, dense_policy="reference")
    return _native_config(wa.config.AdditiveSchwarz(ls))


def reduced_laplacian(B, diag, cfg):
    B = np.ascontiguousarray(B, dtype=np.float64)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    cap = (B.shape[0] + B.shape[1] + 1) ** 2 * 4
    out = np.zeros(cap)
    n = C.c_uint32()
    _lib.check(_lib.lib().wx_testing_reduced_laplacian(B.shape[0], B.shape[1], B.ctypes.data, diag.ctypes.data,
                                                      C.byref(cfg), out.ctypes.data, cap, C.byref(n)))
    return out[:n.value * n.value].reshape(n.value, n.value).copy()


def block_precond(B, diag, cfg) -> Preconditioner:
    B = np.ascontiguousarray(B, dtype=np.float64)
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    h = C.c_void_p()
    _lib.check(_lib.lib().wx_testing_block_precond(B.shape[0], B.shape[1], B.ctypes.data if B.size else None,
                                                  diag.ctypes.data, C.byref(cfg), C.byref(h)))
    return Preconditioner(h)


def dense_block(B, diag):
    nr, nc = B.shape
    A = np.diag(np.asarray(diag, float))
    A[:nr, nr:] = B
    A[nr:, :nr] = B.T
    return A


def dense_exact_schur(Cm, row_diag, col_diag):
    """S = D_c - C' D_r^{-1} C, the reference's test helper (block_elim/schur/tests.rs:18-62)."""
    return np.diag(col_diag) - Cm.T @ np.diag(1.0 / np.asarray(row_diag)) @ Cm


def test_sampled_schur_is_seed_deterministic_and_laplacian(oracle):
    """block_elim/schur/tests.rs:126-165: a floating block's sampled complement is reproducible
    bit for bit from its seed, symmetric, with non-positive off-diagonals, a non-negative
    diagonal and zero row sums."""
    B = np.array([[1.0, 2.0, 3.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    diag = [6.0, 1.0, 1.0, 2.0, 3.0, 3.0]
    a = reduced_laplacian(B, diag, _cfg(seed=12345))
    b = reduced_laplacian(B, diag, _cfg(seed=12345))
    assert a.shape == (3, 3)  # floating: no ground vertex
    assert np.array_equal(a, b)
    assert np.abs(a - a.T).max() <= 1e-12
    off = a - np.diag(np.diag(a))
    assert off.max() <= 1e-12 and np.diag(a).min() >= -1e-12
    assert np.abs(a.sum(axis=1)).max() <= 1e-10
    # the sampled complement is an unbiased estimate: averaged over seeds it approaches the
    # exact complement (a mistake in the sampling weights biases it)
    ex = reduced_laplacian(B, diag, _cfg(exact=True))
    mean = np.mean([reduced_laplacian(B, diag, _cfg(seed=s)) for s in range(4000)], axis=0)
    assert np.abs(mean - ex).max() <= 0.03 * np.abs(ex).max()


def test_sampled_schur_is_exact_on_low_degree_stars():
    """block_elim/schur/tests.rs:167-189: every eliminated star has at most two entries
    including the ground, so sampling is exact; the surplus becomes an explicit ground vertex
    and the augmented matrix is a Laplacian."""
    B = np.array([[1.0, 0.0], [2.0, 3.0], [0.0, 4.0]])
    diag = [3.0, 5.0, 6.0, 3.5, 7.0]
    for seed in (0, 7, 99):
        S = reduced_laplacian(B, diag, _cfg(seed=seed))
        assert S.shape == (3, 3)  # two kept levels + the ground vertex
        expected = dense_exact_schur(B, diag[:3], diag[3:])
        assert np.abs(S[:2, :2] - expected).max() <= 1e-12
        assert np.abs(S.sum(axis=1)).max() < 1e-12


def test_exact_schur_matches_dense_reference():
    """block_elim/schur/tests.rs:84-96: the exact complement (rows eliminated, n_rows >= n_cols)
    equals the dense formula."""
    B = np.array([[1.0, 2.0], [3.0, 0.0], [0.0, 4.0]])
    diag = [5.0, 6.0, 8.0, 7.0, 9.0]
    S = reduced_laplacian(B, diag, _cfg(exact=True))
    expected = dense_exact_schur(B, diag[:3], diag[3:])
    assert np.abs(S[:2, :2] - expected).max() <= 1e-12


def test_grounded_two_block_solve(oracle):
    """block_elim/solver/tests.rs:158-207: a grounded 7 x 7 block (5 eliminated, 2 kept) through
    the sparse path (dense threshold 0, exact Schur): the local solve satisfies A x = r to 1e-9."""
    B = np.zeros((5, 2))
    B[0, 0] = B[1, 1] = 1.0
    diag = [2.0, 3.0, 1.0, 1.0, 1.0, 2.0, 3.0]
    p = block_precond(B, diag, _cfg(exact=True))
    assert set(parse_precond(p.to_bytes())["fac_kind"].tolist()) == {1}
    sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
    r = np.array([0.5, 3.0, -1.25, 0.75, 2.0, 1.0, -2.0])
    x = sch.apply(r)
    assert np.abs(dense_block(B, diag) @ x - r).max() < 1e-9


@pytest.mark.parametrize("shape", [(1, 0), (0, 1)])
def test_trivial_singleton_component_solves_r_over_d(oracle, shape):
    """block_elim/solver/tests.rs:209-236: a live 1 x 1 component solves x = r / d exactly."""
    B = np.zeros(shape)
    p = block_precond(B, [4.0], _cfg(exact=True))
    sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
    assert sch.apply(np.array([2.0]))[0] == 0.5


def test_cover_factor_solves_the_signed_system(oracle):
    """block_elim/factor.rs:182-211: a frustrated (signed) block goes through the Gremban cover's
    factor, and the embed [b, -b] / read-back (y+ - y-)/2 solves the signed system itself."""
    # the 4-cycle has one negative entry: frustrated, so no signature folds it to a Laplacian;
    # with no surplus the block [[2I, B], [B', 2I]] is still definite (eigenvalues 2 +- sqrt 2)
    # and the cover's reduced graph is a 4-cycle, whose elimination is exact at every step
    # (stars of two neighbours), as the reference's exact backend makes its 4 x 4 cover factor
    B = np.array([[1.0, 1.0], [1.0, -1.0]])
    diag = [2.0, 2.0, 2.0, 2.0]
    for cfg in (_cfg(exact=True), _cfg(seed=3)):
        p = block_precond(B, diag, cfg)
        assert set(parse_precond(p.to_bytes())["fac_kind"].tolist()) == {2}
        sch = oracle_schwarz_from_bytes(oracle, p.to_bytes())
        A = dense_block(B, diag)
        for b in (np.array([1.0, 0.5, 0.0, 0.0]), np.array([0.3, -1.0, 2.0, 0.25])):
            x = sch.apply(b)
            assert np.linalg.norm(A @ x - b) < 1e-9


def test_approximate_factor_is_seed_deterministic_and_spd(oracle):
    """Same seeds, same factor bytes; the applied local solve of a larger floating block is a
    symmetric positive semi-definite operator whose null space is the block's own (the all-ones
    direction in folded coordinates), whatever the sampled factor."""
    rng = np.random.default_rng(8)
    nr, nc = 60, 25
    B = np.where(rng.uniform(size=(nr, nc)) < 0.15, rng.uniform(0.5, 2.0, (nr, nc)), 0.0)
    B[np.arange(nr), rng.integers(0, nc, nr)] += 1.0  # every row connected
    diag = np.r_[B.sum(axis=1), B.sum(axis=0)]       # a Laplacian block: floating
    cfg = _cfg(seed=5, chol_seed=9)
    b1 = block_precond(B, diag, cfg).to_bytes()
    assert block_precond(B, diag, cfg).to_bytes() == b1
    assert block_precond(B, diag, _cfg(seed=6, chol_seed=9)).to_bytes() != b1
    sch = oracle_schwarz_from_bytes(oracle, b1)
    n = nr + nc
    M = np.column_stack([sch.apply(e) for e in np.eye(n)])
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    ev = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert ev.min() > -1e-10 * ev.max()
    # the block's null direction (fold: +1 rows, -1 columns) is annihilated
    z = np.r_[np.ones(nr), -np.ones(nc)]
    assert np.abs(M @ z).max() <= 1e-9 * np.abs(M).max() * np.sqrt(n)
    # on the block's range the approximate inverse is within a modest factor of the exact one:
    # the eigenvalues of M A are 0 once (the null direction) and O(1) otherwise
    ev = np.sort(np.linalg.eigvals(M @ dense_block(B, diag)).real)
    assert abs(ev[0]) < 1e-8 and 0.2 < ev[1] and ev[-1] < 5.0
