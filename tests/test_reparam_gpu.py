"""Slope reparametrisation (crates/within/src/solver/reparam.rs:78-276) built on the device against
the host build, and both against what the whitening must satisfy: per level, the weighted Gram of
the rewritten loadings is the identity on the kept coordinates (centred when the effect has an
intercept) and zero elsewhere, and the rank drops are the reference's exact-zero cases
(tests/slopes.rs:77-195).  The device forms the moments in two shifted passes where the host runs
the reference's one-pass Welford update, so the two agree to rounding, not bitwise."""
import ctypes as C

import numpy as np
import pytest

import within_amd as wa
from within_amd import _lib
from within_amd._within import _make_design

from tests._configs import cfg3_data

REL = 1e-12  # device vs host, relative to the size of the whitened loadings


def reparam(effects, w, on_device):
    L = _lib.lib()
    d = _make_design(effects)
    try:
        n = len(effects[0].levels)
        S = sum(len(e.slopes or []) for e in effects)
        nl = [int(np.max(e.levels)) + 1 for e in effects]
        tcap = sum(nl[t] * len(e.slopes or []) ** 2 for t, e in enumerate(effects))
        mcap = sum(nl[t] * len(e.slopes or []) for t, e in enumerate(effects))
        slopes = np.zeros(n * max(S, 1))
        tr = np.zeros(max(tcap, 1))
        mn = np.zeros(max(mcap, 1))
        ucap = 3 * sum(nl[t] * (1 + len(e.slopes or [])) for t, e in enumerate(effects))
        un = np.zeros(max(ucap, 3), np.uint64)
        nu = C.c_uint64()
        wp = None if w is None else np.ascontiguousarray(w, float)
        _lib.check(L.wx_testing_reparam(d, None if wp is None else wp.ctypes.data, 1 if on_device else 0,
                                        slopes.ctypes.data, tr.ctypes.data, tcap, mn.ctypes.data, mcap,
                                        un.ctypes.data, ucap, C.byref(nu)))
    finally:
        L.wx_design_release(d)
    return dict(u=slopes[: n * S].reshape(S, n), transform=tr[:tcap], mean=mn[:mcap],
                unidentified=[tuple(int(x) for x in un[3 * i: 3 * i + 3]) for i in range(nu.value)])


def assert_whitened(levels, u, w, intercept, L, unidentified, term, slope_cols):
    """Per level: sum w u u' = I on the kept coordinates (after centring with an intercept)."""
    w = np.ones(len(levels)) if w is None else np.asarray(w, float)
    pos = w > 0
    v = u.shape[0]
    dropped = {(l, c) for (t, l, c) in unidentified if t == term}
    for l in np.unique(levels[pos]):
        rows = (levels == l) & pos
        U = u[:, rows]
        ww = w[rows]
        G = (U * ww) @ U.T
        kept = sum((l, c) not in dropped for c in slope_cols)
        expect = np.zeros((v, v))
        expect[:kept, :kept] = np.eye(kept)
        assert np.allclose(G, expect, atol=1e-9), (l, G)


def compare(effects, w):
    host = reparam(effects, w, False)
    dev = reparam(effects, w, True)
    assert dev["unidentified"] == host["unidentified"]
    scale = 1.0 + np.abs(host["u"])
    err = np.max(np.abs(dev["u"] - host["u"]) / scale) if host["u"].size else 0.0
    assert err <= REL, err
    assert np.allclose(dev["transform"], host["transform"], rtol=1e-11, atol=1e-12)
    assert np.allclose(dev["mean"], host["mean"], rtol=1e-12, atol=1e-13)
    return host, dev, err


def _panel(nw, weighted, seed=3):
    worker, firm, year, z, _ = cfg3_data(seed=seed, nw=nw, nf=max(nw // 8, 10), k=1)
    w = np.random.default_rng(seed).uniform(0.5, 2.0, len(firm)) if weighted else None
    effects = [wa.Effect(worker, True), wa.Effect(firm, True, [z]), wa.Effect(year, True)]
    return effects, firm, w


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True])
def test_panel_slope_device_matches_host_and_whitens(weighted):
    effects, firm, w = _panel(nw=20_000, weighted=weighted)
    host, dev, _ = compare(effects, w)
    assert_whitened(firm, dev["u"], w, True, None, dev["unidentified"], 1, [1])


@pytest.mark.gpu
def test_three_slopes_with_exact_rank_drops_agree():
    rng = np.random.default_rng(5)
    n, L = 4000, 60
    levels = rng.integers(0, L, n).astype(np.uint32)
    z1 = rng.standard_normal(n)
    z2 = 2.0 * z1 + 1.0                      # collinear with z1 and the intercept everywhere
    z3 = rng.standard_normal(n)
    z3[levels % 7 == 3] = 1.25               # constant inside some levels: dropped there
    w = rng.uniform(0.5, 2.0, n)
    effects = [wa.Effect(levels, True, [z1, z2, z3])]
    host, dev, _ = compare(effects, w)
    drops = dev["unidentified"]
    assert any(c == 3 for (_, l, c) in drops if l % 7 == 3)
    assert all(c != 0 for (_, _, c) in drops)  # the intercept stays identified
    assert_whitened(levels, dev["u"], w, True, L, drops, 0, [1, 2, 3])


@pytest.mark.gpu
def test_slope_only_term_and_zero_weight_garbage():
    rng = np.random.default_rng(9)
    n, L = 3000, 40
    levels = rng.integers(0, L, n).astype(np.uint32)
    levels[levels == 17] = 18                # level 17 unobserved
    z = rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    w[levels == 5] = 0.0                     # level 5 carries only zero weights ...
    z[levels == 5] = 1e150                   # ... and loadings that must not reach the moments
    other = rng.integers(0, 25, n).astype(np.uint32)
    effects = [wa.Effect(levels, False, [z, np.cos(np.arange(n))]), wa.Effect(other, True)]
    host, dev, _ = compare(effects, w)
    drops = dev["unidentified"]
    assert (0, 5, 0) in drops and (0, 5, 1) in drops
    assert (0, 17, 0) in drops and (0, 17, 1) in drops
    assert_whitened(levels, dev["u"], w, False, L, drops, 0, [0, 1])


@pytest.mark.gpu
def test_intercept_of_all_zero_weight_level_is_unidentified():
    rng = np.random.default_rng(2)
    n = 500
    levels = rng.integers(0, 12, n).astype(np.uint32)
    z = rng.standard_normal(n)
    w = np.ones(n)
    w[levels == 4] = 0.0
    host, dev, _ = compare([wa.Effect(levels, True, [z])], w)
    assert (0, 4, 0) in dev["unidentified"] and (0, 4, 1) in dev["unidentified"]


@pytest.mark.gpu
def test_config3_full_size_device_matches_host():
    """The worker-firm panel at its full size (2,000,000 rows, 50,000 firms, firm slope)."""
    effects, firm, w = _panel(nw=400_000, weighted=False)
    host, dev, err = compare(effects, w)
    print(f"config-3 reparametrisation: max relative device-host difference {err:.2e}", flush=True)


def test_host_reparam_whitens_per_level():
    """The host build alone (no GPU): the whitening identity and the exact-zero drops."""
    rng = np.random.default_rng(4)
    n, L = 2000, 30
    levels = rng.integers(0, L, n).astype(np.uint32)
    z1 = rng.standard_normal(n)
    z2 = rng.standard_normal(n)
    z2[levels == 3] = -0.5
    w = rng.uniform(0.5, 2.0, n)
    r = reparam([wa.Effect(levels, True, [z1, z2])], w, False)
    assert r["unidentified"] == [(0, 3, 2)]
    assert_whitened(levels, r["u"], w, True, L, r["unidentified"], 0, [1, 2])
