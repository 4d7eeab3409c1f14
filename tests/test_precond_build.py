"""The product's host preconditioner builder pinned against an independent construction.

`_dense_schwarz_reference` (tests/test_oracle.py) builds the additive Schwarz operator
sum_i R_i' W_i A_i^+ W_i R_i of an intercept-only design directly with numpy/scipy: factor-pair
Gramian blocks over the observed levels, their connected components by scipy's csgraph, and the
partition-of-unity weights 1/sqrt(count) (crates/within/src/domain/factor_pairs.rs:36-194,
domain/cross_tab.rs:25-265).  Here the builder of libwithin_amd.so (wx_precond_build, host only)
must reproduce (a) the same subdomains and weights and (b) with every kept block factored exactly,
the same operator.  A wrong partition-of-unity exponent, a dropped component, a wrong orientation or
a wrong cross-tab would only slow the solver down; these tests catch it directly."""
import numpy as np
import pytest

import within_amd as wa
from oracle.oracle import sparse_design
from tests._helpers import oracle_schwarz_from_bytes, parse_precond
from tests.test_oracle import _dense_schwarz_reference
from within_amd._within import build_preconditioner

EXACT = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=10 ** 6, dense_policy="reference"))


def _designs():
    rng = np.random.default_rng(17)
    out = {}
    n = 1500
    out["two_factor"] = (np.column_stack([rng.integers(0, 60, n), rng.integers(0, 25, n)]), None)
    out["three_factor"] = (np.column_stack([rng.integers(0, 50, n), rng.integers(0, 20, n), rng.integers(0, 6, n)]),
                           None)
    out["weighted"] = (np.column_stack([rng.integers(0, 40, n), rng.integers(0, 30, n), rng.integers(0, 5, n)]),
                       rng.uniform(0.5, 2.0, n))
    # unobserved levels: the codes skip values, so some columns of D are zero
    a = rng.choice(np.r_[0:30:2], n)
    b = rng.choice([0, 1, 2, 5, 9, 11], n)
    out["unobserved_levels"] = (np.column_stack([a, b]), None)
    # disconnected: two blocks of levels that never meet, plus a level seen with one partner only
    h = n // 2
    a = np.r_[rng.integers(0, 20, h), rng.integers(20, 35, n - h)]
    b = np.r_[rng.integers(0, 8, h), rng.integers(8, 14, n - h)]
    c = np.r_[rng.integers(0, 3, h), rng.integers(3, 5, n - h)]
    a[0], b[0] = 35, 14  # a pair seen once: its own component in (A, B)
    out["disconnected"] = (np.column_stack([a, b, c]), rng.uniform(0.3, 3.0, n))
    # a zero weight removes an observation from every cross-tab
    codes = np.column_stack([rng.integers(0, 30, 800), rng.integers(0, 12, 800)])
    w = rng.uniform(0.5, 1.5, 800)
    w[:40] = 0.0
    out["zero_weights"] = (codes, w)
    return {k: (np.asfortranarray(c.astype(np.uint32)), w) for k, (c, w) in out.items()}


DESIGNS = _designs()


def _components(p):
    """(sorted globals, PoU weights) per component of the exported preconditioner."""
    out = []
    ss = p["slot_start"]
    for c in range(len(ss) - 1):
        g = p["glob"][ss[c]:ss[c + 1]].astype(np.int64)
        pw = p["pw"][ss[c]:ss[c + 1]]
        o = np.argsort(g)
        out.append((tuple(g[o]), pw[o]))
    return out


def _slot_map(p):
    """(component, global DOF) -> (PoU weight, fold factor, 1/D_e)."""
    ss = p["slot_start"]
    out = {}
    for c in range(len(ss) - 1):
        for s in range(int(ss[c]), int(ss[c + 1])):
            out[(c, int(p["glob"][s]))] = (float(p["pw"][s]), float(p["fold"][s]), float(p["inv_de"][s]))
    return out


def _cross_tab(p):
    """(global eliminated DOF, global kept DOF) -> C entry, over every component."""
    ss = p["slot_start"]
    out = {}
    for c in range(len(ss) - 1):
        s0, ne = int(ss[c]), int(p["ne"][c])
        for e in range(ne):
            s = s0 + e
            for q in range(int(p["c_ptr"][s]), int(p["c_ptr"][s + 1])):
                kslot = s0 + ne + int(p["c_idx"][q])
                out[(int(p["glob"][s]), int(p["glob"][kslot]))] = float(p["c_val"][q])
    return out


@pytest.mark.parametrize("case", list(DESIGNS))
def test_subdomains_and_weights_match_independent_build(case):
    codes, w = DESIGNS[case]
    pre = build_preconditioner(codes, w, EXACT)
    p = parse_precond(pre.to_bytes())
    M, comps = _dense_schwarz_reference(codes, w)
    ref = {tuple(sorted(ii)) for ii in comps}
    got = _components(p)
    assert {g for g, _ in got} == ref
    count = np.zeros(M.shape[0])
    for ii in comps:
        count[ii] += 1
    for g, pw in got:
        np.testing.assert_allclose(pw, 1.0 / np.sqrt(count[list(g)]), rtol=1e-15)
    # the eliminated block of a plain pair is its larger side (sddm.rs:86-106)
    n_levels = np.array([int(codes[:, t].max()) + 1 for t in range(codes.shape[1])])
    offs = np.concatenate([[0], np.cumsum(n_levels)])
    ss = p["slot_start"]
    for c in range(len(ss) - 1):
        g = p["glob"][ss[c]:ss[c + 1]].astype(np.int64)
        ne, nk = int(p["ne"][c]), int(p["nk"][c])
        te = np.searchsorted(offs, g[:ne], side="right") - 1
        tk = np.searchsorted(offs, g[ne:], side="right") - 1
        assert len(set(te)) == 1 and len(set(tk)) <= 1
        assert ne >= nk


@pytest.mark.parametrize("case", list(DESIGNS))
def test_exact_schwarz_apply_matches_independent_operator(oracle, case):
    codes, w = DESIGNS[case]
    pre = build_preconditioner(codes, w, EXACT)
    M, _ = _dense_schwarz_reference(codes, w)
    sch = oracle_schwarz_from_bytes(oracle, pre.to_bytes())
    rng = np.random.default_rng(5)
    X = rng.standard_normal((M.shape[0], 3))
    got = np.column_stack([sch.apply(X[:, j]) for j in range(3)])
    ref = M @ X
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_default_factors_differ_from_exact_only_inside_kept_blocks(oracle):
    """The default local solver (sampled Schur + approximate Cholesky above the dense threshold)
    keeps the same subdomains, weights, orientation and cross-tabs; only the kept-block factor
    differs.  Its apply is a symmetric positive semi-definite operator with the exact operator's
    null space (the levels no subdomain covers)."""
    codes, w = DESIGNS["weighted"]
    ex = parse_precond(build_preconditioner(codes, w, EXACT).to_bytes())
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    ac_bytes = build_preconditioner(codes, w, cfg).to_bytes()
    ac = parse_precond(ac_bytes)
    for f in ("slot_start", "ne", "nk"):
        assert np.array_equal(ex[f], ac[f]), f
    # kept slots of a factored block are stored in elimination order: compare per global DOF
    assert _slot_map(ex) == _slot_map(ac)
    assert _cross_tab(ex) == _cross_tab(ac)
    assert set(ac["fac_kind"].tolist()) == {1}
    sch = oracle_schwarz_from_bytes(oracle, ac_bytes)
    n = ac["n"]
    Mac = np.column_stack([sch.apply(e) for e in np.eye(n)])
    assert np.abs(Mac - Mac.T).max() <= 1e-12 * np.abs(Mac).max()
    assert np.linalg.eigvalsh(0.5 * (Mac + Mac.T)).min() > -1e-10 * np.abs(Mac).max()
    D = sparse_design(codes).toarray()
    unseen = np.flatnonzero((D.T @ (w[:, None] * D)).diagonal() == 0)
    assert np.all(Mac[unseen] == 0.0)
