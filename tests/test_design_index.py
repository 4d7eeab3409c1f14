"""The design's index structures, built by the library, against the reference's known answers
(crates/within/src/domain.rs:355-518): the locality sort's permutation, the sorted flags, the
dominant term and the term-major layout.  Host only (wx_testing_design_index)."""
import ctypes as C

import numpy as np
import pytest

import within_amd as wa
from within_amd import _lib
from within_amd._within import _make_design


def index(design):
    L = _lib.lib()
    d = _make_design(design)
    try:
        T = design.shape[1] if isinstance(design, np.ndarray) else len(design)
        n = design.shape[0] if isinstance(design, np.ndarray) else len(design[0].levels)
        nl = np.zeros(T, np.uint64)
        off = np.zeros(T, np.uint64)
        srt = np.zeros(T, np.uint8)
        dom = C.c_int64()
        permuted = C.c_int32()
        perm = np.zeros(max(n, 1), np.uint32)
        _lib.check(L.wx_testing_design_index(d, T, nl.ctypes.data, off.ctypes.data, srt.ctypes.data, C.byref(dom),
                                             C.byref(permuted), perm.ctypes.data))
    finally:
        L.wx_design_release(d)
    return dict(n_levels=nl.tolist(), offset=off.tolist(), sorted=[bool(v) for v in srt], dominant=dom.value,
                obs_perm=perm[:n].tolist() if permuted.value else None)


def codes(*cols):
    return np.asfortranarray(np.column_stack(cols).astype(np.uint32))


def test_from_frame_sorts_owned_unsorted_dominant():
    # domain.rs:433-447: factor 0 (3 levels) dominates and is unsorted
    r = index(codes([2, 0, 1, 0], [0, 0, 1, 1]))
    assert r["dominant"] == 0
    assert r["obs_perm"] == [1, 3, 2, 0]
    assert r["sorted"] == [True, False]  # factor 1's permuted column [0,1,1,0] is not sorted


def test_rescan_marks_nested_factor_sorted_after_permutation():
    # domain.rs:449-457
    col0 = np.array([3, 0, 2, 1])
    r = index(codes(col0, col0 // 2))
    assert r["obs_perm"] is not None
    assert r["sorted"] == [True, True]


def test_from_frame_keeps_sorted_input():
    # domain.rs:459-466
    r = index(codes([0, 0, 1, 2], [1, 0, 1, 0]))
    assert r["obs_perm"] is None
    assert r["sorted"] == [True, False]


def test_dominant_is_the_last_argmax_of_n_dofs():
    # ties go to the last index (Rust max_by_key, domain.rs:226): both terms have 4 levels
    r = index(codes([3, 1, 0, 2, 1], [0, 3, 2, 1, 0]))
    assert r["dominant"] == 1
    assert r["obs_perm"] == [0, 4, 3, 2, 1]  # stable argsort of [0,3,2,1,0]


def test_new_lays_out_slope_terms_term_major():
    # domain.rs:482-518: term 0 = [intercept, z0, z1] over 2 levels, term 1 = intercept over 3
    # levels, term 2 = one slope over 2 levels
    f0 = np.array([0, 0, 1, 1], np.uint32)
    f1 = np.array([0, 2, 1, 0], np.uint32)
    z0, z1 = np.array([1.0, 2, 3, 4]), np.array([5.0, 6, 7, 8])
    r = index([wa.Effect(f0, True, [z0, z1]), wa.Effect(f1, True), wa.Effect(f0, False, [z1])])
    assert r["offset"] == [0, 6, 9]
    assert r["n_levels"] == [2, 3, 2]
    assert r["obs_perm"] is None  # the dominant term (term 0, 6 dofs) is sorted


def _xorshift_keys(n, span):
    state = 0x2545F4914F6CDD1D
    mask = (1 << 64) - 1
    out = np.empty(n, np.uint64)
    for i in range(n):
        state ^= (state << 13) & mask
        state ^= state >> 7
        state ^= (state << 17) & mask
        out[i] = state % span
    return out.astype(np.uint32)


@pytest.mark.parametrize("n_obs,key_span", [(1, 1), (997, 1), (997, 16), (997, 996), (997, 997), (997, 998),
                                            (997, 50_000), (4096, 8), (4096, 4096)])
def test_locality_sort_is_a_stable_argsort(n_obs, key_span):
    """domain.rs:355-392: both branches of stable_argsort (counting sort when n >= L, comparison
    sort otherwise) give the stable ascending permutation.  A one-factor design exercises the
    library's sort on the same keys (n_levels = max key + 1)."""
    key = _xorshift_keys(n_obs, key_span)
    r = index(codes(key))
    expected = np.argsort(key, kind="stable").tolist()
    if expected == list(range(n_obs)):
        assert r["obs_perm"] is None and r["sorted"] == [True]
    else:
        assert r["obs_perm"] == expected and r["sorted"] == [True]


def test_locality_sort_of_a_large_design_is_the_stable_argsort_of_the_dominant_term():
    rng = np.random.default_rng(9)
    n = 200_000
    a = rng.integers(0, 5000, n)
    b = rng.integers(0, 300, n)
    c = rng.integers(0, 12, n)
    r = index(codes(b, a, c))
    assert r["dominant"] == 1
    assert r["obs_perm"] == np.argsort(a, kind="stable").tolist()
    assert r["sorted"] == [False, True, False]
    assert r["offset"] == [0, 300, 5300]
