"""The reference's serialized preconditioner (postcard wire v13) against its own fixtures.

The byte files under tests/golden/reference_wire/ are the reference's test fixtures, kept as data:
preconditioner_v13.postcard and preconditioner_v12.postcard (crates/within/tests/fixtures/,
driven by crates/within/tests/wire_format_fixture.rs:29-96) and the two fuzzer crash inputs of
crates/within/tests/precond_deser_robustness.rs:10-50.  The v13 fixture is the default
preconditioner of a 6-row design with a frustrated (f-slope, g) pair
(wire_format_fixture.rs:14-27); its within-level data -- subdomains, partition weights, cross
blocks in SDDM coordinates, eliminated-diagonal reciprocals and the coordinate map -- is compared
bit for bit with this library's host build of the same design.  Host only."""
import os

import numpy as np
import pytest

import within_amd as wa
from within_amd import wire_v13
from within_amd._within import build_preconditioner

from tests._helpers import parse_precond

HERE = os.path.join(os.path.dirname(__file__), "golden", "reference_wire")


def fixture(name):
    with open(os.path.join(HERE, name), "rb") as f:
        return f.read()


def fixture_design():
    # wire_format_fixture.rs:14-27
    f = np.array([0, 0, 0, 1, 1, 1], np.uint32)
    g = np.array([0, 1, 2, 0, 1, 2], np.uint32)
    z = np.array([-2.0, 1.0, 1.0, -1.0, -1.0, 2.0])
    return [wa.Effect(f, True, [z]), wa.Effect(g, True)]


def test_v13_fixture_decodes_to_the_fixture_problem():
    ref = wire_v13.decode(fixture("preconditioner_v13.postcard"))
    assert ref.variant == "Additive" and ref.n_dofs == 7
    a, b = ref.subdomains
    # (f intercept, g) and (f slope, g): g's three levels eliminated, f's two kept
    assert a.global_indices.tolist() == [4, 5, 6, 0, 1]
    assert b.global_indices.tolist() == [4, 5, 6, 2, 3]
    assert a.uniform and b.uniform            # slope designs keep uniform weights (factor_pairs.rs:82-88)
    assert (a.n_eliminated, a.n_kept) == (3, 2) and (b.n_eliminated, b.n_kept) == (3, 2)
    assert np.array_equal(a.c.dense(), np.ones((3, 2)))
    assert np.array_equal(a.inv_diag_elim, [0.5, 0.5, 0.5])
    assert a.reduced_factor == "Direct" and a.grounding == "Floating" and a.coordinates is None
    assert b.reduced_factor == "Cover" and b.cover_m == 2           # the frustrated pair's cover
    assert np.allclose(b.coordinates, [np.sqrt(6) / 4] * 3 + [-1.0, -1.0], rtol=0, atol=1e-15)
    # C and C' agree
    assert np.array_equal(b.ct.dense(), b.c.dense().T)


def test_host_build_reproduces_the_reference_within_level_data_bitwise():
    ref = wire_v13.decode(fixture("preconditioner_v13.postcard"))
    ours = parse_precond(bytes(build_preconditioner(fixture_design()).__reduce__()[1][0]))
    assert ours["n"] == ref.n_dofs
    assert len(ours["ne"]) == len(ref.subdomains)
    for i, s in enumerate(ref.subdomains):
        a, b = int(ours["slot_start"][i]), int(ours["slot_start"][i + 1])
        ne, nk = int(ours["ne"][i]), int(ours["nk"][i])
        assert (ne, nk) == (s.n_eliminated, s.n_kept)
        assert ours["glob"][a:b].tolist() == s.global_indices.tolist()
        assert np.array_equal(ours["pw"][a:b], s.partition_weights)
        cs, ce = int(ours["c_ptr"][a]), int(ours["c_ptr"][a + ne])
        assert np.array_equal(ours["c_ptr"][a:a + ne + 1] - cs, s.c.indptr)
        assert np.array_equal(ours["c_idx"][cs:ce], s.c.indices)
        assert np.array_equal(ours["c_val"][cs:ce], s.c.data)          # bit for bit
        assert np.array_equal(ours["inv_de"][a:a + ne], s.inv_diag_elim)
        # the coordinate map: Canonical negates the kept block (sddm.rs:50-75)
        fold = ours["fold"][a:b]
        expect = s.coordinates if s.coordinates is not None else np.r_[np.ones(ne), -np.ones(nk)]
        assert np.array_equal(fold, expect)


def test_pre_bump_v12_fixture_no_longer_decodes():
    # wire_format_fixture.rs:94-96
    with pytest.raises(wire_v13.WireError, match="v12"):
        wire_v13.decode(fixture("preconditioner_v12.postcard"))


@pytest.mark.parametrize("name", ["precond_deser_crash_ndofs_span.bin", "precond_deser_crash_nested_cover.bin"])
def test_fuzzer_crash_inputs_are_rejected(name):
    # precond_deser_robustness.rs:22-31
    with pytest.raises(wire_v13.WireError):
        wire_v13.decode(fixture(name))
    with pytest.raises(ValueError):
        wa.Preconditioner(fixture(name))


def test_malformed_inputs_are_rejected():
    # precond_deser_robustness.rs:33-50
    crash = fixture("precond_deser_crash_ndofs_span.bin")
    for b in [b"", b"\x00", b"\xff" * 8, b"\xff" * 64, crash[: len(crash) // 2], crash[:1]]:
        with pytest.raises(wire_v13.WireError):
            wire_v13.decode(b)
        with pytest.raises(ValueError):
            wa.Preconditioner(b)


def test_every_truncation_of_the_v13_fixture_is_rejected():
    b = fixture("preconditioner_v13.postcard")
    for m in range(len(b)):
        with pytest.raises(wire_v13.WireError):
            wire_v13.decode(b[:m])


def test_n_dofs_below_the_covered_span_is_rejected():
    b = bytearray(fixture("preconditioner_v13.postcard"))
    assert b[-1] == 7
    b[-1] = 6                                   # the subdomains index DOF 6
    with pytest.raises(wire_v13.WireError, match="n_dofs"):
        wire_v13.decode(bytes(b))
    b[-1] = 9                                   # uncovered DOFs are allowed
    assert wire_v13.decode(bytes(b)).n_dofs == 9


def test_additive_payload_is_reported_not_loaded():
    with pytest.raises(ValueError, match="approx-chol"):
        wa.Preconditioner(fixture("preconditioner_v13.postcard"))


def test_reference_diagonal_payload_loads_and_applies():
    inv = np.array([0.5, 0.25, 2.0, 1.0])
    b = bytes([1, len(inv)]) + inv.astype("<f8").tobytes()   # Variant::Diagonal{inv_diag}
    ref = wire_v13.decode(b)
    assert ref.variant == "Diagonal" and np.array_equal(ref.inv_diag, inv)
    p = wa.Preconditioner(b)
    assert (p.nrows, p.ncols) == (4, 4)
    out = parse_precond(bytes(p.__reduce__()[1][0]))
    assert out["kind"] == 0 and np.array_equal(out["inv_diag"], inv)
