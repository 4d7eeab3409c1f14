"""Host (CPU) tests of the additive Schwarz builder over designs with slope channels.

The preconditioner is built by the product's host code (wx_precond_build, no GPU), decoded
from its bytes and applied by the oracle.  With every kept block factored exactly, each local
solve must equal F M^{+/-1} F for the component's Gramian G folded to M = F G F by the stored
coordinate map F with the diagonal finalised as in crates/within/src/domain/factor_pairs/
sddm.rs:281-336 (raised to the magnitude row sums, replaced by them when the surplus is
round-off).  The Gramian comes from numpy with the slope whitened per level as the
reparametrisation does (crates/within/src/solver/reparam.rs; one slope: centre, normalise).
"""
import numpy as np
import pytest

import within_amd as wa
from within_amd._within import build_preconditioner
from oracle.oracle import sparse_design
from tests._helpers import oracle_schwarz_from_bytes, parse_precond, whiten


def dense_reference(p, G):
    """sum_c R_c' diag(pw) F (M_c)^{+/-1} F diag(pw) R_c from the decoded component table."""
    n = G.shape[0]
    ref = np.zeros((n, n))
    eps = np.finfo(float).eps
    for c in range(len(p["ne"])):
        s0, s1 = int(p["slot_start"][c]), int(p["slot_start"][c + 1])
        ne = int(p["ne"][c])
        idx = p["glob"][s0:s1].astype(np.int64)
        F, pw = p["fold"][s0:s1], p["pw"][s0:s1]
        M = F[:, None] * G[np.ix_(idx, idx)] * F[None, :]
        off = M.copy()
        np.fill_diagonal(off, 0.0)
        rs = np.abs(off).sum(1)
        d = np.maximum(np.diag(M), rs)
        sur = d - rs
        floating = sur.sum() <= 4 * eps * np.sqrt(len(d)) * d.sum()
        np.fill_diagonal(M, rs if floating else d)
        assert floating == (p["gauge"][c] == 0)
        # the stored elimination data are this matrix's
        np.testing.assert_allclose(p["inv_de"][s0:s0 + ne], 1.0 / np.diag(M)[:ne], rtol=1e-12)
        if floating:
            ev, V = np.linalg.eigh(M)
            keep = ev > ev.max() * 1e-11
            Mi = (V[:, keep] / ev[keep]) @ V[:, keep].T
        else:
            Mi = np.linalg.inv(M)
        ref[np.ix_(idx, idx)] += (pw[:, None] * F[:, None]) * Mi * (F[None, :] * pw[None, :])
    return ref


@pytest.mark.parametrize("weighted", [False, True])
def test_slope_design_exact_local_solves(oracle, weighted):
    rng = np.random.default_rng(21 + weighted)
    n = 2500
    a = rng.integers(0, 60, n).astype(np.uint32)
    b = rng.integers(0, 25, n).astype(np.uint32)
    c = rng.integers(0, 4, n).astype(np.uint32)
    z = rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    design = [wa.Effect(a, True), wa.Effect(b, True, [z]), wa.Effect(c, True)]
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=10 ** 6))
    pre = build_preconditioner(design, w, cfg)
    p = parse_precond(pre.to_bytes())
    assert set(p["fac_kind"].tolist()) == {0}  # every kept block exact and dense
    assert (p["gauge"] == 1).any()  # slope pairs retain surplus (grounded) or are signed
    assert (p["pw"] == 1.0).all()  # uniform partition weights with a slope channel
    wv = np.ones(n) if w is None else w
    D = sparse_design(np.column_stack([a, b, c]), effects=[(True, []), (True, [whiten(b, z, wv)]), (True, [])])
    G = (D.T @ (wv[:, None] * D.toarray()))
    ref = dense_reference(p, G)
    sch = oracle_schwarz_from_bytes(oracle, pre.to_bytes())
    for _ in range(3):
        r = rng.standard_normal(G.shape[0])
        np.testing.assert_allclose(sch.apply(r), ref @ r, rtol=1e-8, atol=1e-10)


def test_slope_design_sparse_factors_are_spd(oracle):
    # approximate Cholesky everywhere (dense_threshold=0): signed kept blocks go through the
    # Gremban cover (factor kind 2); the apply stays symmetric positive semi-definite
    rng = np.random.default_rng(31)
    n = 3000
    a = rng.integers(0, 80, n).astype(np.uint32)
    b = rng.integers(0, 30, n).astype(np.uint32)
    z = rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    design = [wa.Effect(a, True), wa.Effect(b, True, [z])]
    cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    pre = build_preconditioner(design, w, cfg)
    p = parse_precond(pre.to_bytes())
    assert 2 in set(p["fac_kind"].tolist())  # a frustrated slope component uses the cover
    sch = oracle_schwarz_from_bytes(oracle, pre.to_bytes())
    N = int(p["n"])
    M = np.column_stack([sch.apply(e) for e in np.eye(N)])
    np.testing.assert_allclose(M, M.T, rtol=1e-9, atol=1e-11)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > -1e-9


def test_preconditioner_bytes_round_trip_general():
    rng = np.random.default_rng(41)
    n = 1000
    a = rng.integers(0, 40, n).astype(np.uint32)
    b = rng.integers(0, 15, n).astype(np.uint32)
    design = [wa.Effect(a, True), wa.Effect(b, True, [rng.standard_normal(n)])]
    for thr in (0, 10 ** 6):
        cfg = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=thr, dense_policy="reference"))
        pre = build_preconditioner(design, None, cfg)
        b1 = pre.to_bytes()
        assert wa.Preconditioner(b1).to_bytes() == b1
        # corrupt a cover permutation / a fold factor: rejected, never a fault
        bad = bytearray(b1)
        bad[-5] ^= 0xFF
        with pytest.raises(ValueError):
            wa.Preconditioner(bytes(bad))
