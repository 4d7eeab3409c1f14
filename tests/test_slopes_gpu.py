"""Varying slopes through the device solver: coefficients in the caller's parametrisation, the
fitted values, and the rank-drop (`unidentified`) report, restating the known answers of the
reference's crates/within/tests/slopes.rs:19-340 and tests/metamorphic.rs:136-160.  The per-level
OLS closed form and the exact-zero rule for unidentified slots pin x itself, which the
gauge-free demeaned checks elsewhere cannot see."""
import numpy as np
import pytest

import within_amd as wa

pytestmark = pytest.mark.gpu
TOL = 1e-6


def close(actual, expected):
    actual, expected = np.asarray(actual, float), np.asarray(expected, float)
    assert np.all(np.abs(actual - expected) <= TOL * (1.0 + np.abs(expected))), (actual, expected)


def per_level_intercept_slope(levels, z, y, w, L):
    """Per-level weighted OLS of y on [1, z] (slopes.rs:19-45)."""
    w = np.ones(len(y)) if w is None else np.asarray(w)
    sw = np.bincount(levels, w, L)
    sz = np.bincount(levels, w * z, L)
    szz = np.bincount(levels, w * z * z, L)
    sy = np.bincount(levels, w * y, L)
    szy = np.bincount(levels, w * z * y, L)
    with np.errstate(divide="ignore", invalid="ignore"):
        b = (sw * szy - sz * sy) / (sw * szz - sz * sz)
        a = (sy - b * sz) / sw
    return a, b


def synthetic_y(n):
    return np.sin(np.arange(n) * 0.9 - 1.0) * 2.0 + 0.5


def drops(r):
    return [(d.term, d.level, d.column) for d in r.unidentified]


def solve_with(levels, z, intercept, w, y, pc=None):
    levels = np.asarray(levels, np.uint32)
    zs = [np.asarray(v, float) for v in (z if isinstance(z, list) else [z])]
    s = wa.Solver([wa.Effect(levels, intercept, zs)], weights=None if w is None else np.asarray(w, float),
                  preconditioner=pc)
    return s.solve(np.asarray(y, float))


PCS = [None, wa.PreconditionerConfig.Diagonal]


@pytest.mark.parametrize("pc", PCS)
def test_single_slope_recovers_per_level_ols_and_fitted_values(gpu, pc):
    levels = np.array([2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1], np.uint32)
    z = np.array([0.5, -1.2, 3.3, 1.5, 0.7, -2.1, 2.5, 1.9, 0.4, -0.5, -1.0, 1.1])
    y = synthetic_y(len(levels))
    r = solve_with(levels, z, True, None, y, pc)
    assert r.converged and drops(r) == []
    a, b = per_level_intercept_slope(levels, z, y, None, 3)
    close(r.x[:3], a)
    close(r.x[3:6], b)
    close(y - r.demeaned, a[levels] + b[levels] * z)


def test_slope_only_term_recovers_per_level_projection(gpu):
    levels = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2], np.uint32)
    z = np.array([1.0, 2.0, -1.5, 0.5, 3.0, 1.0, -2.0, 0.7, 1.3])
    y = synthetic_y(len(levels))
    r = solve_with(levels, z, False, None, y)
    assert r.converged and drops(r) == [] and len(r.x) == 3
    b = np.bincount(levels, z * y) / np.bincount(levels, z * z)
    close(r.x, b)
    close(y - r.demeaned, r.x[levels] * z)


def test_rank_drops_report_deterministically_with_exact_zeros(gpu):
    # level 1 never occurs; level 2's slope is constant, so only its slope drops
    levels = np.array([0, 0, 0, 2, 2, 2, 3, 3, 3], np.uint32)
    z = np.array([1.0, 2.0, -1.5, 2.5, 2.5, 2.5, -2.0, 0.7, 1.3])
    y = synthetic_y(len(levels))
    r = solve_with(levels, z, True, None, y)
    assert r.converged
    assert drops(r) == [(0, 1, 0), (0, 1, 1), (0, 2, 1)]
    for slot in (1, 4 + 1, 4 + 2):
        assert r.x[slot] == 0.0
    assert np.all(np.isfinite(r.x))
    close(r.x[2], (y[3] + y[4] + y[5]) / 3.0)
    a, b = per_level_intercept_slope(levels, z, y, None, 4)
    for l in (0, 3):
        close(r.x[l], a[l])
        close(r.x[4 + l], b[l])
    d = solve_with(levels, z, True, None, y, wa.PreconditionerConfig.Diagonal)
    assert drops(d) == drops(r)
    close(d.x, r.x)
    # the unidentified directions index the layout (metamorphic.rs:136-160)
    for u in r.unidentified:
        assert r.x[r.layout.index(u.term, u.level, u.column)] == 0.0


def test_weighted_solve_matches_closed_form_and_masks_zero_weight_rows(gpu):
    levels = np.array([0, 1, 0, 1, 0, 1, 2, 2, 2], np.uint32)
    z = np.array([0.5, -1.2, 3.3, 1.5, 0.7, -2.1, 5.0, 5.0, 7.0])
    w = np.array([1.0, 0.5, 2.0, 1.5, 3.0, 1.0, 1.0, 1.0, 0.0])
    y = synthetic_y(len(levels))
    r = solve_with(levels, z, True, w, y)
    assert r.converged
    assert drops(r) == [(0, 2, 1)]
    assert r.x[3 + 2] == 0.0
    a, b = per_level_intercept_slope(levels, z, y, w, 3)
    for l in range(2):
        close(r.x[l], a[l])
        close(r.x[3 + l], b[l])
    close(r.x[2], (y[6] + y[7]) / 2.0)


def test_zero_weight_garbage_does_not_poison_identification(gpu):
    levels = np.array([0, 0, 0, 1, 1, 1], np.uint32)
    z = np.array([1.0, 2.0, 3.0, 5.0, 7.0, 1e300])
    w = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.0])
    y = synthetic_y(len(levels))
    r = solve_with(levels, z, True, w, y)
    assert drops(r) == []
    a, b = per_level_intercept_slope(levels[:5], z[:5], y[:5], w[:5], 2)
    for l in range(2):
        close(r.x[l], a[l])
        close(r.x[2 + l], b[l])
    d = solve_with(levels, z, True, w, y, wa.PreconditionerConfig.Diagonal)
    close(d.x, r.x)


def test_batch_solve_shares_unidentified_and_back_transforms_each_rhs(gpu):
    levels = np.array([0, 0, 0, 1, 1, 1], np.uint32)
    z = np.array([1.0, 2.0, 3.0, 2.5, 2.5, 2.5])
    y1 = synthetic_y(len(levels))
    y2 = y1 * -1.5 + 0.3
    s = wa.Solver([wa.Effect(levels, True, [z])])
    batch = s.solve_batch(np.asfortranarray(np.column_stack([y1, y2])))
    assert drops(batch) == [(0, 1, 1)]
    s1, s2 = s.solve(y1), s.solve(y2)
    assert np.array_equal(batch.x[:, 0], s1.x) and np.array_equal(batch.x[:, 1], s2.x)


def test_three_slopes_solve_bounded_with_exact_rank_drops(gpu):
    n = 24
    i = np.arange(n)
    levels = (i % 3).astype(np.uint32)
    z1 = np.sin(i * 0.37) * 2.0
    z2 = 0.95 * z1 + 0.1 * np.cos(i * 0.9 + 0.2)
    z3 = np.where(i % 3 == 1, 2.0 * z2, -0.8 * z1 + 0.15 * np.sin(i * 1.7) + 0.1)
    y = synthetic_y(n)
    r = solve_with(levels, [z1, z2, z3], True, None, y)
    assert r.converged and r.iterations <= 30
    # the pivot keeps the larger-variance z3, so z2's column of level 1 reads exactly 0
    assert drops(r) == [(0, 1, 2)]
    assert r.x[2 * 3 + 1] == 0.0
    for l in range(3):
        m = levels == l
        for col in (np.ones(n), z1, z2, z3):
            assert abs(np.sum(r.demeaned[m] * col[m])) < 1e-6
    fitted = r.x[levels] + r.x[3 + levels] * z1 + r.x[6 + levels] * z2 + r.x[9 + levels] * z3
    close(y - r.demeaned, fitted)


def test_panel_slope_unidentified_slots_are_zero_at_scale(gpu):
    """A worker-firm panel with a firm slope: firms seen in one observation (or with a constant
    covariate) have an unidentified slope, reported in ascending (level, column) order and held at
    exactly 0; every identified coefficient satisfies the per-firm normal equations."""
    rng = np.random.default_rng(12)
    nw, T, nf = 3000, 3, 900
    worker = np.repeat(np.arange(nw), T).astype(np.uint32)
    firm = rng.integers(0, nf, nw * T).astype(np.uint32)
    firm[::97] = nf + np.arange(len(firm[::97]))  # singleton firms
    z = rng.standard_normal(nw * T)
    z[firm == 5] = 1.25                            # a constant covariate within one firm
    y = rng.standard_normal(nw * T)
    s = wa.Solver([wa.Effect(worker, True), wa.Effect(firm, True, [z])])
    r = s.solve(y, wa.LsmrOptions(tol=1e-12))
    assert r.converged
    d = drops(r)
    assert d == sorted(d)
    L = int(firm.max()) + 1
    counts = np.bincount(firm, minlength=L)
    expect = sorted({(1, int(f), 1) for f in np.flatnonzero(counts == 1)} | {(1, 5, 1)})
    assert d == expect
    for u in r.unidentified:
        assert r.x[r.layout.index(u.term, u.level, u.column)] == 0.0
    # per-firm normal equations of the slope column hold on the demeaned residual
    g = np.bincount(firm, r.demeaned * z, L)
    assert np.abs(g).max() < 1e-8 * np.abs(y).sum()
