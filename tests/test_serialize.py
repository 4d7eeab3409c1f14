"""Preconditioner bytes are untrusted input (they arrive through pickle): every malformed byte
string must be rejected with a ValueError, never crash or read out of bounds -- the property
the reference pins in crates/within/tests/precond_deser_robustness.rs:22-50.  Host-only: the
preconditioners here are built and decoded without a GPU."""
import numpy as np
import pytest

import within_amd as wa
from within_amd._within import build_preconditioner


def _blobs():
    rng = np.random.default_rng(91)
    n = 3000
    a = rng.integers(0, 120, n).astype(np.uint32)
    b = (rng.pareto(1.2, n) * 4).astype(np.int64).clip(0, 59).astype(np.uint32)
    c = rng.integers(0, 6, n).astype(np.uint32)
    w = rng.uniform(0.5, 2.0, n)
    sparse = wa.config.AdditiveSchwarz(wa.config.LocalSolverConfig(dense_threshold=0, dense_policy="reference"))
    out = {
        "dense": build_preconditioner(np.asfortranarray(np.column_stack([a, b, c])), weights=w),
        "sparse": build_preconditioner(np.asfortranarray(np.column_stack([a, b, c])), weights=w,
                                       preconditioner=sparse),
        # a slope channel: signed components solved through Gremban covers
        "cover": build_preconditioner([wa.Effect(a, True), wa.Effect(b, True, [rng.standard_normal(n)])],
                                      weights=w, preconditioner=sparse),
        "diagonal": build_preconditioner(np.asfortranarray(np.column_stack([a, b])),
                                         preconditioner=wa.PreconditionerConfig.Diagonal),
    }
    return {k: p.to_bytes() for k, p in out.items()}


BLOBS = None


def blobs():
    global BLOBS
    if BLOBS is None:
        BLOBS = _blobs()
    return BLOBS


def rejected(data: bytes) -> bool:
    try:
        wa.Preconditioner(data)
    except ValueError as e:
        assert "deserialize" in str(e)
        return True
    return False


@pytest.mark.parametrize("kind", ["dense", "sparse", "cover", "diagonal"])
def test_round_trip_is_byte_identical(kind):
    b = blobs()[kind]
    p = wa.Preconditioner(b)
    assert p.to_bytes() == b
    assert p.nrows == p.ncols


@pytest.mark.parametrize("kind", ["dense", "sparse", "cover", "diagonal"])
def test_every_truncation_is_rejected(kind):
    b = blobs()[kind]
    cuts = sorted(set(range(0, len(b), max(1, len(b) // 400))) | {1, len(b) // 2, len(b) - 1})
    for cut in cuts:
        assert rejected(b[:cut]), cut


@pytest.mark.parametrize("kind", ["dense", "sparse", "cover"])
def test_corrupted_bytes_are_rejected_or_decode_consistently(kind):
    """Random byte overwrites: a structural field that breaks the format is rejected; a value
    that keeps it valid (a coefficient, say) decodes to a preconditioner that re-encodes to
    exactly the corrupted bytes."""
    b = blobs()[kind]
    rng = np.random.default_rng(5)
    n_rejected = 0
    for _ in range(300):
        m = bytearray(b)
        for off in rng.integers(0, len(b), rng.integers(1, 4)):
            m[off] = int(rng.integers(0, 256))
        m = bytes(m)
        if rejected(m):
            n_rejected += 1
        else:
            assert wa.Preconditioner(m).to_bytes() == m
    assert n_rejected > 0


def test_structural_corruptions_are_rejected():
    b = blobs()["sparse"]
    for bad in [b"", b"\x00", b"\xff" * 8, b"\xff" * 64, b[: len(b) // 2], b[:1], b + b"\x00"]:
        assert rejected(bad), len(bad)
    # n (the DOF count, right after the magic, version and kind) below the covered slots' span
    # (the reference's first fuzz find)
    head = bytearray(b)
    n_off = 4 + 4 + 4
    head[n_off:n_off + 8] = (1).to_bytes(8, "little")
    assert rejected(bytes(head))
