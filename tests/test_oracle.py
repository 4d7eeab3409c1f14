"""Pins the CPU oracle against the reference's known answers and closed forms (CPU only)."""
import json
import os

import numpy as np
import pytest

from oracle.oracle import OracleSchwarz, exact_demean, normal_eq_residual, sparse_design

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))
SMALL = np.load(os.path.join(HERE, "golden", "small_projections.npz"))


def xorshift_keys(n, span, state=0x2545F4914F6CDD1D):
    mask = (1 << 64) - 1
    out = []
    for _ in range(n):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        out.append(state % span)
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("n,L,span", [(0, 0, 1), (0, 4, 4), (1, 1, 1), (997, 1, 1), (997, 16, 16), (997, 996, 996),
                                      (997, 997, 997), (997, 998, 998), (997, 50_000, 50_000), (4096, 4096, 8),
                                      (4096, 4096, 4096)])
def test_stable_argsort_branches(oracle, n, L, span):
    # crates/within/src/domain.rs:356-392
    key = xorshift_keys(n, span)
    got = oracle.stable_argsort(key, L)
    assert np.array_equal(got, np.argsort(key, kind="stable").astype(np.uint32))


def test_design_layout_known_answers(oracle):
    ka = KNOWN["locality_sort"]
    d = oracle.design(np.array(ka["codes"]).T)
    assert d.dominant == 0 and not d.sorted[0]
    perm = oracle.stable_argsort(d.codes[:, d.dominant], int(d.n_levels[d.dominant]))
    assert perm.tolist() == ka["obs_perm"]
    dp = oracle.permuted(d, perm)
    assert dp.sorted.tolist() == ka["sorted"]
    assert dp.codes.T.tolist() == ka["permuted_codes"]
    nested = oracle.permuted(oracle.design(np.array(KNOWN["nested_rescan"]["codes"]).T),
                             np.argsort(np.array(KNOWN["nested_rescan"]["codes"][0]), kind="stable"))
    assert nested.sorted.tolist() == KNOWN["nested_rescan"]["sorted"]
    ks = oracle.design(np.array(KNOWN["keeps_sorted"]["codes"]).T)
    assert ks.sorted.tolist() == KNOWN["keeps_sorted"]["sorted"]
    lay = KNOWN["layout"]
    d = oracle.design(np.array(lay["levels"]).T, ncols=lay["ncols"], colload=[-1, 0, 1, -1, 2],
                      slopes=np.ones((4, 3)))
    assert d.offset.tolist() == lay["offsets"] and d.n_dofs == lay["n_dofs"]


def test_dominant_is_last_argmax(oracle):
    # ties go to the LAST maximal term (Rust max_by_key, domain.rs:226)
    codes = np.array([[0, 1, 2, 0], [2, 1, 0, 1]]).T
    assert oracle.design(codes).dominant == 1


def test_gather_scatter_known_answers(oracle):
    ka = KNOWN["design_ops"]
    d = oracle.design(np.array(ka["codes"]).T)
    assert d.n_dofs == ka["n_dofs"]
    assert oracle.gather(d, np.array(ka["x"])).tolist() == ka["Dx"]
    assert oracle.scatter(d, np.array(ka["r"])).tolist() == ka["Dtr"]


@pytest.mark.parametrize("weighted", [False, True])
def test_gather_scatter_match_scipy_and_adjoint(oracle, weighted):
    rng = np.random.default_rng(7)
    n = 2000
    codes = np.column_stack([rng.integers(0, 50, n), rng.integers(0, 9, n), rng.integers(0, 3, n)])
    z = rng.standard_normal(n)
    d = oracle.design(codes, ncols=[1, 2, 1], colload=[-1, -1, 0, 0], slopes=z.reshape(-1, 1))
    D = sparse_design(codes, [(True, []), (True, [z]), (False, [z])])
    sw = np.sqrt(rng.uniform(0.5, 2, n)) if weighted else None
    x = rng.standard_normal(d.n_dofs)
    u = rng.standard_normal(n)
    s = np.ones(n) if sw is None else sw
    np.testing.assert_allclose(oracle.gather(d, x, sw), s * (D @ x), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.scatter(d, u, sw), D.T @ (s * u), rtol=1e-12, atol=1e-12)
    lhs = oracle.gather(d, x, sw) @ u
    rhs = x @ oracle.scatter(d, u, sw)
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))


def test_diagonal_closed_form(oracle):
    rng = np.random.default_rng(3)
    n = 500
    codes = np.column_stack([rng.integers(0, 30, n), rng.integers(0, 6, n)])
    w = rng.uniform(0.1, 3, n)
    d = oracle.design(codes)
    inv = oracle.build_diagonal(d, w)
    diag = np.concatenate([np.bincount(codes[:, 0], weights=w, minlength=30),
                           np.bincount(codes[:, 1], weights=w, minlength=6)])
    np.testing.assert_allclose(inv, np.where(diag > 0, 1 / np.where(diag > 0, diag, 1), 0), rtol=1e-14)


def test_vec_norm(oracle):
    for v, expect in KNOWN["lsmr"]["vec_norm"]["cases"]:
        assert oracle.vec_norm(np.array(v, dtype=float)) == expect
    assert np.isfinite(oracle.vec_norm(np.array([1e300, 1e300])))
    assert np.isnan(oracle.vec_norm(np.array([np.nan])))
    assert np.isnan(oracle.vec_norm(np.array([0.0, np.nan])))


def test_lsmr_reference_fixtures(oracle):
    F = KNOWN["lsmr"]
    od = F["overdetermined"]
    for prec, data in [("none", None), ("diag", np.array(od["diag"])), ("identity", None)]:
        rc, x, res = oracle.lsmr_dense(np.array(od["A"], float), np.array(od["b"], float), od["tol"], 100, prec, data)
        assert rc == 0 and res.converged
        assert np.linalg.norm(x - od["x"]) < 1e-6
    for key in ["beta_zero", "inconsistent_breakdown"]:
        c = F[key]
        for prec in ["none", "identity"]:
            rc, x, res = oracle.lsmr_dense(np.array(c["A"], float), np.array(c["b"], float), c["tol"], 100, prec)
            assert rc == 0 and res.converged and res.iterations == c["iterations"]
            from oracle.oracle import STOP_NAMES
            assert STOP_NAMES[res.stop_reason] == c["stop"]
            np.testing.assert_allclose(x, c["x"], atol=1e-12)
            if "residual_norm" in c:
                assert abs(res.residual_norm - c["residual_norm"]) < 1e-12
    c = F["alpha1_zero"]
    rc, x, res = oracle.lsmr_dense(np.array(c["A"], float), np.array(c["b"], float), c["tol"], 100)
    assert res.iterations == 0 and x.tolist() == c["x"] and res.residual_norm == c["residual_norm"]
    c = F["underdetermined"]
    rc, x, res = oracle.lsmr_dense(np.array(c["A"], float), np.array(c["b"], float), c["tol"], 100)
    np.testing.assert_allclose(x, c["x"], atol=1e-10)
    # maxiter = 1 -> MaxIterations (lsmr/tests/solve.rs:187-196)
    A = np.random.default_rng(0).standard_normal((30, 12))
    rc, x, res = oracle.lsmr_dense(A, np.ones(30), 1e-14, 1)
    assert res.iterations == 1 and not res.converged and res.stop_reason == 5


def test_lsmr_matches_scipy(oracle):
    import scipy.sparse.linalg as spl
    rng = np.random.default_rng(11)
    A = rng.standard_normal((60, 25))
    b = rng.standard_normal(60)
    _, x, res = oracle.lsmr_dense(A, b, 1e-12, 500)
    xs = spl.lsmr(A, b, atol=1e-14, btol=1e-14, maxiter=500)[0]
    np.testing.assert_allclose(x, xs, rtol=1e-8, atol=1e-10)
    xe = np.linalg.lstsq(A, b, rcond=None)[0]
    np.testing.assert_allclose(x, xe, rtol=1e-8, atol=1e-10)


def test_mlsmr_indefinite_preconditioner_rejected(oracle):
    # bidiag.rs:99-111: <v, M v> < -sqrt(eps)|v||p| is an error
    A = np.eye(3)
    rc, _, _ = oracle.lsmr_dense(A, np.ones(3), 1e-10, 10, "dense", -np.eye(3))
    assert rc == -2


def test_exact_demean_two_closed_forms():
    # the dense pseudo-inverse projection agrees with scipy LSQR driven to precision
    import scipy.sparse.linalg as spl
    codes, y = SMALL["two_factor__codes"], SMALL["two_factor__y"]
    D = sparse_design(codes)
    ex = exact_demean(D, y)
    for j in range(y.shape[1]):
        x = spl.lsqr(D, y[:, j], atol=1e-15, btol=1e-15, iter_lim=50000)[0]
        np.testing.assert_allclose(ex[:, j], y[:, j] - D @ x, atol=1e-10)
        assert normal_eq_residual(codes, y[:, j], ex[:, j])[0] < 1e-12


@pytest.mark.parametrize("case", ["two_factor", "three_factor_weighted", "disconnected"])
@pytest.mark.parametrize("prec", ["off", "diagonal"])
def test_oracle_solve_matches_golden(oracle, case, prec):
    codes, y = SMALL[f"{case}__codes"], SMALL[f"{case}__y"]
    w = SMALL[f"{case}__w"] if f"{case}__w" in SMALL.files else None
    d = oracle.design(codes)
    X, Dm, res = oracle.solve_batch(d, y, w, prec=prec, tol=1e-12, maxiter=5000)
    assert all(r.converged for r in res)
    np.testing.assert_allclose(Dm, SMALL[f"{case}__demeaned"], atol=1e-8)


def _dense_schwarz_reference(codes, w):
    """sum over factor-pair components of R' W A^+ W R built densely with numpy."""
    D = sparse_design(codes).toarray()
    W = np.ones(len(codes)) if w is None else w
    G = D.T @ (W[:, None] * D)
    n_levels = [int(codes[:, t].max()) + 1 for t in range(codes.shape[1])]
    offs = np.concatenate([[0], np.cumsum(n_levels)])
    comps = []
    for a in range(codes.shape[1]):
        for b in range(a + 1, codes.shape[1]):
            idx = [i for i in range(offs[a], offs[a + 1]) if G[i, i] > 0] + \
                  [i for i in range(offs[b], offs[b + 1]) if G[i, i] > 0]
            sub = G[np.ix_(idx, idx)].copy()
            # connected components of the pair graph
            import scipy.sparse.csgraph as cg
            ncomp, lab = cg.connected_components(np.abs(sub) > 0, directed=False)
            for c in range(ncomp):
                ii = [idx[q] for q in range(len(idx)) if lab[q] == c]
                comps.append(ii)
    count = np.zeros(G.shape[0])
    for ii in comps:
        count[ii] += 1
    M = np.zeros_like(G)
    for ii in comps:
        pw = 1 / np.sqrt(count[ii])
        Ai = G[np.ix_(ii, ii)]
        M[np.ix_(ii, ii)] += (pw[:, None] * np.linalg.pinv(Ai, rcond=1e-12)) * pw[None, :]
    return M, comps


def _export_exact_schwarz(oracle, codes, w):
    """Exact (dense) Schwarz in the oracle's exported layout, built with numpy only."""
    D = sparse_design(codes).toarray()
    W = np.ones(len(codes)) if w is None else w
    G = D.T @ (W[:, None] * D)
    _, comps = _dense_schwarz_reference(codes, w)
    count = np.zeros(G.shape[0])
    for ii in comps:
        count[ii] += 1
    A = {k: [] for k in ["slot_start", "ne", "nk", "glob", "pw", "inv_de", "c_ptr", "c_idx", "c_val", "fac_kind",
                         "fac_off", "fac_val", "piv_off", "piv", "fold", "gauge"]}
    slot, fo, po, cnnz = 0, 0, 0, 0
    A["slot_start"].append(0)
    n_levels0 = int(codes[:, 0].max()) + 1
    for ii in comps:
        # split by factor (the pair's two blocks), eliminate the larger
        blocks = {}
        for g in ii:
            t = int(np.searchsorted(np.cumsum([int(codes[:, q].max()) + 1 for q in range(codes.shape[1])]), g,
                                    side="right"))
            blocks.setdefault(t, []).append(g)
        keys = sorted(blocks)
        if len(keys) == 1:
            E, K = blocks[keys[0]], []
        else:
            r, c = blocks[keys[0]], blocks[keys[1]]
            E, K = (c, r) if len(c) > len(r) else (r, c)
        ne, nk = len(E), len(K)
        A["ne"].append(ne)
        A["nk"].append(nk)
        A["glob"] += E + K
        A["pw"] += [1 / np.sqrt(count[g]) for g in E + K]
        A["fold"] += [1.0] * ne + [-1.0] * nk  # plain pair: the kept block is negated
        A["gauge"].append(0)
        A["inv_de"] += [1 / G[g, g] for g in E] + [0.0] * nk
        for e in E:
            A["c_ptr"].append(cnnz)
            for kk, g in enumerate(K):
                if G[e, g] != 0:
                    A["c_idx"].append(kk)
                    A["c_val"].append(G[e, g])  # cross-tab C >= 0 of the folded Laplacian [D -C; -C' D']
                    cnnz += 1
        for _ in K:
            A["c_ptr"].append(cnnz)
        S = G[np.ix_(K, K)] - G[np.ix_(K, E)] @ np.diag([1 / G[e, e] for e in E]) @ G[np.ix_(E, K)]
        # LDL' of the grounded Laplacian (last pivot 0): dense unit-lower L
        Lm = np.eye(nk)
        dvec = np.zeros(nk)
        Sw = S.copy()
        for q in range(nk - 1):
            dvec[q] = Sw[q, q]
            Lm[q + 1:, q] = Sw[q + 1:, q] / dvec[q]
            Sw[q + 1:, q + 1:] -= np.outer(Lm[q + 1:, q], Sw[q, q + 1:])
        A["fac_kind"].append(0)
        A["fac_off"].append(fo)
        A["fac_val"] += Lm.ravel().tolist()
        fo += nk * nk
        A["piv_off"].append(po)
        A["piv"] += dvec.tolist()
        po += nk
        slot += ne + nk
        A["slot_start"].append(slot)
    A["c_ptr"].append(cnnz)
    return OracleSchwarz(oracle, G.shape[0], {k: np.array(v) for k, v in A.items()}), n_levels0


@pytest.mark.parametrize("weighted", [False, True])
def test_oracle_schwarz_apply_equals_dense_pinv_sum(oracle, weighted):
    rng = np.random.default_rng(5)
    n = 300
    codes = np.column_stack([rng.integers(0, 25, n), rng.integers(0, 6, n), rng.integers(0, 3, n)]).astype(np.uint32)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    M, _ = _dense_schwarz_reference(codes, w)
    sch, _ = _export_exact_schwarz(oracle, codes, w)
    for _ in range(3):
        r = rng.standard_normal(M.shape[0])
        np.testing.assert_allclose(sch.apply(r), M @ r, rtol=1e-9, atol=1e-9)


def test_oracle_schwarz_solve_matches_golden(oracle):
    codes, y, w = (SMALL["three_factor_weighted__codes"], SMALL["three_factor_weighted__y"],
                   SMALL["three_factor_weighted__w"])
    sch, _ = _export_exact_schwarz(oracle, codes, w)
    d = oracle.design(codes)
    X, Dm, res = oracle.solve_batch(d, y, w, prec="schwarz", schwarz=sch, tol=1e-12)
    assert all(r.converged for r in res)
    np.testing.assert_allclose(Dm, SMALL["three_factor_weighted__demeaned"], atol=1e-8)


# ---------------------------------------------------------------------------
# general (signed) components: fold, gauge none, explicit inverse and Gremban cover
# ---------------------------------------------------------------------------

def _one_component_schwarz(oracle, M, ne, fold, kind, fac):
    """Oracle layout of a single component whose SDDM-coordinate matrix is M (eliminated block
    first, both diagonal blocks diagonal), folded by `fold`, with the kept-block factor `fac`."""
    nk = M.shape[0] - ne
    C = -M[:ne, ne:]
    A = dict(slot_start=[0, ne + nk], ne=[ne], nk=[nk], glob=list(range(ne + nk)), pw=[1.0] * (ne + nk),
             inv_de=[1 / M[e, e] for e in range(ne)] + [0.0] * nk, fold=list(fold), gauge=[1],
             c_ptr=[], c_idx=[], c_val=[], fac_kind=[kind], fac_off=[0], piv_off=[0])
    nnz = 0
    for e in range(ne):
        A["c_ptr"].append(nnz)
        for k in range(nk):
            if C[e, k] != 0:
                A["c_idx"].append(k)
                A["c_val"].append(C[e, k])
                nnz += 1
    A["c_ptr"] += [nnz] * (nk + 1)
    A.update(fac)
    return OracleSchwarz(oracle, ne + nk, {k: np.array(v) for k, v in A.items()})


def _sparse_ldl_rows(S, order):
    """LDL' of S permuted to `order` (no pivoting) as strictly-lower CSR rows; a pivot that is
    round-off of a singular (floating) system becomes 0."""
    P = S[np.ix_(order, order)].copy()
    n = len(order)
    L = np.eye(n)
    d = np.zeros(n)
    scale = np.abs(np.diag(S)).max()
    for q in range(n):
        d[q] = P[q, q] - L[q, :q] ** 2 @ d[:q]
        if abs(d[q]) < 1e-12 * scale:
            d[q] = 0.0
        for i in range(q + 1, n):
            v = P[i, q] - (L[i, :q] * L[q, :q]) @ d[:q]
            L[i, q] = v / d[q] if d[q] != 0 else 0.0
    ptr, idx, val = [0], [], []
    for i in range(n):
        for j in range(i):
            if L[i, j] != 0:
                idx.append(j)
                val.append(L[i, j])
        ptr.append(len(idx))
    return ptr, idx, val, d


def test_oracle_general_component_explicit_inverse(oracle):
    # slope channel pair (term 0 intercept, term 1 slope) of a small weighted design: for any
    # coordinate map F, fold -> exact kept inverse of M = F G F -> unfold gives G^-1
    rng = np.random.default_rng(11)
    n = 400
    a, b = rng.integers(0, 12, n), rng.integers(0, 5, n)
    z, w = rng.standard_normal(n), rng.uniform(0.5, 2.0, n)
    ne, nk = 12, 5
    G = np.zeros((ne + nk, ne + nk))
    np.add.at(G, (a, a), w)
    np.add.at(G, (ne + b, ne + b), w * z * z)
    np.add.at(G, (a, ne + b), w * z)
    np.add.at(G, (ne + b, a), w * z)
    assert np.linalg.eigvalsh(G).min() > 0
    F = rng.choice([-1.0, 1.0], ne + nk) * rng.uniform(0.5, 2.0, ne + nk)
    M = F[:, None] * G * F[None, :]
    De, C = np.diag(M)[:ne], -M[:ne, ne:]
    S = np.diag(np.diag(M)[ne:]) - C.T @ np.diag(1 / De) @ C
    sch = _one_component_schwarz(oracle, M, ne, F, 2, dict(fac_val=np.linalg.inv(S).ravel(), piv=[0.0]))
    Gi = np.linalg.inv(G)
    for _ in range(3):
        r = rng.standard_normal(ne + nk)
        np.testing.assert_allclose(sch.apply(r), Gi @ r, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("surplus", [True, False])
def test_oracle_cover_solve_equals_inverse(oracle, surplus):
    # frustrated SDD bipartite component; the reduced solve goes through an LDL' of the
    # Gremban cover's Schur complement in a random elimination order (block_elim/solver.rs:221-277,
    # factor.rs:93-118).  Without surplus the cover is a singular Laplacian (zero last pivot).
    rng = np.random.default_rng(12 + surplus)
    ne, nk = 9, 4
    C = rng.uniform(0.2, 1.0, (ne, nk)) * rng.choice([-1.0, 1.0], (ne, nk))
    C[rng.uniform(size=(ne, nk)) < 0.3] = 0.0
    for e in range(ne):  # keep every eliminated level connected
        if not C[e].any():
            C[e, rng.integers(nk)] = 0.7
    sur = rng.uniform(0.1, 0.5, ne + nk) if surplus else np.zeros(ne + nk)
    d = np.concatenate([np.abs(C).sum(1), np.abs(C).sum(0)]) + sur
    M = np.diag(d)
    M[:ne, ne:] = -C
    M[ne:, :ne] = -C.T
    assert np.linalg.eigvalsh(M).min() > 1e-9  # frustrated: definite even without surplus
    # cover: positive entries join equal sheets, negative ones opposite sheets
    Cc = np.zeros((2 * ne, 2 * nk))
    for e in range(ne):
        for k in range(nk):
            v = C[e, k]
            if v > 0:
                Cc[e, k] = Cc[ne + e, nk + k] = v
            elif v < 0:
                Cc[e, nk + k] = Cc[ne + e, k] = -v
    Dc = np.concatenate([d[:ne], d[:ne], d[ne:], d[ne:]])
    Sc = np.diag(Dc[2 * ne:]) - Cc.T @ np.diag(1 / Dc[:2 * ne]) @ Cc
    order = rng.permutation(2 * nk)
    ptr, idx, val, piv = _sparse_ldl_rows(Sc, order)
    F = rng.choice([-1.0, 1.0], ne + nk) * rng.uniform(0.5, 2.0, ne + nk)
    G = M / F[:, None] / F[None, :]  # original coordinates: M = F G F
    fac = dict(fac_val=val, fac_ptr=ptr, fac_idx=idx, piv=piv, cover_off=[0], cover_perm=order)
    sch = _one_component_schwarz(oracle, M, ne, F, 3, fac)
    Gi = np.linalg.inv(G)
    for _ in range(3):
        r = rng.standard_normal(ne + nk)
        np.testing.assert_allclose(sch.apply(r), Gi @ r, rtol=1e-9, atol=1e-11)


# ---------------------------------------------------------------------------
# windowed reorthogonalisation (crates/schwarz-precond/src/lsmr/tests/solve.rs:330-432)
# ---------------------------------------------------------------------------

def _vandermonde(rows=30, cols=12):
    x = np.arange(rows) / max(rows - 1, 1)
    A = np.vander(x, cols, increasing=True)
    b = np.log1p(x)
    return A, b


def _normal_residual(A, x, b):
    return np.linalg.norm(A.T @ (b - A @ x))


def test_oracle_reorth_unpreconditioned_converges(oracle):
    A, b = _vandermonde()
    rc0, x0, r0 = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=30)
    rc, x, r = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=30, local_size=10)
    assert rc == 0 and r.converged
    assert not r0.converged or r.iterations <= r0.iterations
    assert _normal_residual(A, x, b) < 1e-6


def test_oracle_reorth_preconditioned_converges(oracle):
    A, b = _vandermonde()
    jac = 1.0 / (A * A).sum(0)
    rc, x, r = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=30, prec="diag", prec_data=jac, local_size=10)
    assert rc == 0 and r.converged
    assert _normal_residual(A, x, b) < 1e-6


@pytest.mark.parametrize("window", [1, 12, 13])
def test_oracle_reorth_window_boundaries(oracle, window):
    A, b = _vandermonde()
    rc, x, r = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=200, local_size=window)
    assert rc == 0 and r.converged
    assert _normal_residual(A, x, b) < 1e-6


def test_oracle_reorth_identity_preconditioner_matches_unpreconditioned(oracle):
    # the M-weighted ring with M = I is the Euclidean one (lsmr.rs: lsmr == mlsmr(M = I))
    A, b = _vandermonde()
    _, x1, r1 = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=30, local_size=10)
    _, x2, r2 = oracle.lsmr_dense(A, b, tol=1e-9, maxiter=30, prec="identity", local_size=10)
    assert r1.iterations == r2.iterations
    # cond(A) ~ 1e10 leaves x ill-determined; the fitted values are not
    np.testing.assert_allclose(A @ x1, A @ x2, rtol=1e-9, atol=1e-9)
