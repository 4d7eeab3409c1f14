"""LSMR's early exits and breakdowns through the device solver, on fixed-effects designs built so
the arithmetic is exact (the reference's cases use toy operators:
crates/schwarz-precond/src/lsmr/tests/breakdown.rs:8-72, 219-243)."""
import struct

import numpy as np
import pytest

import within_amd as wa
from within_amd._within import build_preconditioner

pytestmark = pytest.mark.gpu

PRECONDS = {"off": wa.PreconditionerConfig.Off, "diagonal": wa.PreconditionerConfig.Diagonal,
            "additive": wa.PreconditionerConfig.Additive}


@pytest.mark.parametrize("prec", list(PRECONDS))
def test_alpha1_zero_early_exit(gpu, prec):
    """D'y = 0 with y != 0 (breakdown.rs:40-72): no iteration, x = 0 and the demeaned output is
    y itself."""
    # a full 4 x 3 grid of (a, b) cells, two rows per cell; y = u_a v_b with sum(u) = sum(v) = 0
    a = np.repeat(np.arange(4), 6) % 4
    b = np.tile(np.repeat(np.arange(3), 2), 4)
    codes = np.asfortranarray(np.column_stack([a, b]).astype(np.uint32))
    u = np.array([1.0, -1.0, 2.0, -2.0])
    v = np.array([1.0, 0.0, -1.0])
    y = u[a] * v[b]
    assert np.all(np.bincount(a, weights=y) == 0) and np.all(np.bincount(b, weights=y) == 0)
    r = wa.solve(codes, y, preconditioner=PRECONDS[prec])
    assert r.converged and r.iterations == 0
    assert r.stop_reason == "InitialNormalEquationResidualZero"
    assert np.all(r.x == 0.0)
    assert np.array_equal(r.demeaned, y)
    assert r.residual == 0.0


@pytest.mark.parametrize("prec", ["off", "diagonal"])
def test_beta_zero_breakdown_is_a_converged_residual_stop(gpu, prec):
    """y in the range of D along one coefficient (breakdown.rs:8-34): u_2 = A v_1 - alpha_1 u_1
    collapses to exactly zero, which the solver reports as a converged ResidualTolerance stop
    after one iteration with the exact coefficients."""
    codes = np.array([0, 0, 0, 0, 1, 1, 2, 2, 2], dtype=np.uint32)[:, None]
    y = np.array([5.0, 5, 5, 5, 0, 0, 0, 0, 0])
    r = wa.solve(np.asfortranarray(codes), y, preconditioner=PRECONDS[prec],
                 options=wa.LsmrOptions(tol=1e-12))
    assert r.converged and r.iterations == 1
    assert r.stop_reason == "ResidualTolerance"
    np.testing.assert_allclose(r.x, [5.0, 0.0, 0.0], atol=1e-12)
    assert np.all(r.demeaned == 0.0)


def test_indefinite_preconditioner_is_rejected(gpu):
    """M^{-1} = -diag(D'D)^{-1} (breakdown.rs:219-243): the first <v, p~> is -||p~||_M^2, so the
    solve must fail with the reference's InvalidInput (a ValueError), not return a result."""
    rng = np.random.default_rng(3)
    n = 2000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 40, n), rng.integers(0, 7, n)]).astype(np.uint32))
    y = rng.standard_normal(n)
    p = build_preconditioner(codes, None, wa.PreconditionerConfig.Diagonal)
    raw = bytearray(p.to_bytes())
    magic, ver, kind, nd, ln = struct.unpack_from("<4sIIQQ", raw, 0)
    assert magic == b"WXPC" and nd == ln == 47
    diag = np.frombuffer(bytes(raw[28:28 + 8 * ln]), dtype="<f8")
    assert np.all(diag > 0)
    raw[28:28 + 8 * ln] = (-diag).astype("<f8").tobytes()
    neg = wa.Preconditioner(bytes(raw))
    s = wa.Solver(codes, preconditioner=neg)
    with pytest.raises(ValueError, match="not positive definite") as e:
        s.solve(y)
    assert e.value.variant == "SolveError::InvalidInput"
    # the same design with the genuine preconditioner solves
    assert wa.Solver(codes, preconditioner=p).solve(y).converged
