"""Many independent problems in one call (solve_many).  Unpacked shapes (one run per problem on
one or several streams, with or without captured iteration graphs) run the same sequence of
kernels per problem as the Solver's own solve_batch, and the extra iterations a group or the
lag-one poll runs past convergence are gated off, so their output is bitwise equal.  Packed
shapes run problems of one shape as one block-diagonal batch: the same algorithm per problem
with the norm and dot-product partial sums grouped differently, so they agree to rounding."""
import numpy as np
import pytest

import within_amd as wa
from tests._configs import cfg4_problems

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def problems():
    out = []
    for codes, Y in cfg4_problems(seed=41, n_problems=20, lo=3.3, hi=5.0):
        out.append((wa.Solver(codes), Y))
    rng = np.random.default_rng(7)
    # weighted, and one problem wider than one device batch (two chunks)
    n = 30_000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 1500, n), rng.integers(0, 90, n)]).astype(np.uint32))
    out.append((wa.Solver(codes, weights=np.exp(rng.normal(0, 0.5, n))), np.asfortranarray(rng.standard_normal((n, 70)))))
    # a slope term (host back-transform of x) under the Schwarz preconditioner
    n = 20_000
    a, b = rng.integers(0, 800, n).astype(np.uint32), rng.integers(0, 60, n).astype(np.uint32)
    z = rng.standard_normal(n)
    out.append((wa.Solver([wa.Effect(a, True), wa.Effect(b, True, [z])]), np.asfortranarray(rng.standard_normal((n, 3)))))
    # pairs that pack together: weighted (Schwarz), Diagonal, Off; k 3 and 4 share a width class
    for pc in (None, wa.PreconditionerConfig.Diagonal, wa.PreconditionerConfig.Off):
        for q, k in enumerate((3, 4)):
            n = 8_000 + 3_000 * q
            codes = np.asfortranarray(np.column_stack([rng.integers(0, n // 20, n), rng.integers(0, 40, n),
                                                       rng.integers(0, 7, n)]).astype(np.uint32))
            w = np.exp(rng.normal(0, 0.4, n)) if pc is None else None
            s = wa.Solver(codes, weights=w) if pc is None else wa.Solver(codes, preconditioner=pc)
            Y = np.asfortranarray(rng.standard_normal((n, k)))
            if pc is wa.PreconditionerConfig.Off:
                EXACT[len(out)] = exact_demeaned(codes, Y)
            out.append((s, Y))
    return out


EXACT = {}  # problem index -> exact projection (the unpreconditioned problems)


def exact_demeaned(codes, Y):
    """Y - D x with x = argmin ||Y - D x|| by dense least squares (small designs)."""
    cols = []
    for t in range(codes.shape[1]):
        L = int(codes[:, t].max()) + 1
        Dt = np.zeros((codes.shape[0], L))
        Dt[np.arange(codes.shape[0]), codes[:, t]] = 1.0
        cols.append(Dt)
    D = np.hstack(cols)
    x = np.linalg.lstsq(D, Y, rcond=None)[0]
    return Y - D @ x


@pytest.fixture(scope="module")
def singles(problems):
    opts = wa.LsmrOptions(tol=1e-8)
    return [s.solve_batch(Y, opts) for s, Y in problems]


# (streams, iterations per graph, graph, pack[, work-item tile of the packs])
SHAPES = [(1, 1, False, False), (4, 1, True, False), (8, 2, True, False), (16, 4, True, False),
          (1, 1, False, True), (4, 1, True, True), (4, 4, True, True), (4, 2, True, True, 32),
          (4, 2, True, True, 256)]


@pytest.mark.parametrize("shape", SHAPES)
def test_many_matches_individual_solves(gpu, problems, singles, shape):
    print(f"[many] shape {shape}", flush=True)
    wa.set_many_shape(*shape)
    try:
        res = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], wa.LsmrOptions(tol=1e-8))
    finally:
        wa.set_many_shape(0)
    got = wa.last_many_shape()
    assert (got["streams"], got["iterations_per_graph"], got["graph"], got["pack"]) == shape[:4]
    assert got["tile"] == (shape[4] if len(shape) > 4 else 64)
    packed = shape[3]
    worst_d = worst_x = worst_off = worst_plain = 0.0
    diter = n_off = n_cols = 0
    for p, (r, ref) in enumerate(zip(res, singles)):
        plain = problems[p][0].preconditioner is None
        assert all(r.converged), p
        assert r.demeaned.shape == ref.demeaned.shape and r.x.shape == ref.x.shape
        if not packed:
            assert r.iterations == ref.iterations, p
            np.testing.assert_allclose(r.true_residual, ref.true_residual, rtol=1e-9, atol=1e-15)
        for j in range(r.demeaned.shape[1]):
            n_cols += 1
            dd = rel(r.demeaned[:, j], ref.demeaned[:, j])
            if not packed:
                assert dd <= 1e-13, (p, j)
            if plain:
                # plain LSMR (~40 steps, no reorthogonalisation) amplifies rounding
                # differences chaotically: both answers are held to the exact projection
                ex = EXACT[p][:, j]
                worst_plain = max(worst_plain, rel(r.demeaned[:, j], ex), rel(ref.demeaned[:, j], ex))
            elif r.iterations[j] == ref.iterations[j]:
                worst_d = max(worst_d, dd)
                worst_x = max(worst_x, rel(r.x[:, j], ref.x[:, j]))
            else:
                # a borderline stopping test decided one step apart: both are converged solutions
                n_off += 1
                diter = max(diter, abs(r.iterations[j] - ref.iterations[j]))
                worst_off = max(worst_off, dd)
        assert max(r.true_residual) < 1e-5, p
        pd = max(rel(r.demeaned[:, j], ref.demeaned[:, j]) for j in range(r.demeaned.shape[1]))
        if pd > 1e-11:
            print(f"  problem {p}: n={r.demeaned.shape[0]} k={r.demeaned.shape[1]} "
                  f"precond={repr(problems[p][0].preconditioner)[:30]} "
                  f"iterations {r.iterations} vs {ref.iterations}, demeaned {pd:.2e}")
    print(f"shape {shape}: same-iteration columns demeaned {worst_d:.2e}, x {worst_x:.2e} (unpreconditioned "
          f"{worst_plain:.2e}); {n_off} of {n_cols} columns stop {diter} step(s) apart, demeaned {worst_off:.2e}")
    if packed:
        # the same iterates to rounding (partial sums grouped differently)
        assert worst_d <= 1e-10 and worst_x <= 1e-9 and worst_plain <= 1e-6
        assert diter <= 1 and n_off <= max(2, n_cols // 10) and worst_off <= 1e-6
    else:
        assert n_off == 0 and worst_d <= 1e-13 and worst_x <= 1e-12 and worst_plain <= 1e-6


def test_many_autotuned_mixed_precision(gpu, problems):
    opts = wa.LsmrOptions(tol=1e-8, precision="mixed")
    res = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], opts)
    shape = wa.last_many_shape()
    # a second call reuses the cached shape
    res2 = wa.solve_many([s for s, _ in problems], [Y for _, Y in problems], opts)
    assert wa.last_many_shape() == shape
    # packed runs agree with solve_batch to rounding (to the tolerance where a borderline
    # stopping test decides one step apart), unpacked ones bitwise
    for p, ((s, Y), r, r2) in enumerate(zip(problems, res, res2)):
        ref = s.solve_batch(Y, opts)
        assert all(r.converged) and all(r2.converged), p
        for j in range(Y.shape[1]):
            if not shape["pack"]:
                assert r.iterations[j] == ref.iterations[j]
                bound = 1e-13
            else:
                assert abs(r.iterations[j] - ref.iterations[j]) <= 1
                plain = s.preconditioner is None
                bound = 1e-6 if (plain or r.iterations[j] != ref.iterations[j]) else 1e-10
            assert rel(r.demeaned[:, j], ref.demeaned[:, j]) <= bound
            assert rel(r2.demeaned[:, j], r.demeaned[:, j]) <= 1e-13 # repeatable
    print(f"autotuned shape {shape}")


@pytest.mark.parametrize("shape", [(4, 2, True, False), (4, 2, True, True)])
def test_many_graph_replay_reads_the_current_tolerance(gpu, problems, shape):
    """A later call on the same Solvers replays the iteration graphs captured by an earlier one:
    it must still stop at its own tolerance (the tolerance lives in the per-column device
    state, not in the captured launches)."""
    sub = problems[:4] + problems[-6:]
    solvers, Ys = [s for s, _ in sub], [Y for _, Y in sub]
    wa.set_many_shape(*shape)
    try:
        loose = wa.solve_many(solvers, Ys, wa.LsmrOptions(tol=1e-6))
        tight = wa.solve_many(solvers, Ys, wa.LsmrOptions(tol=1e-12))
    finally:
        wa.set_many_shape(0)
    more = 0
    for p, ((s, Y), rl, rt) in enumerate(zip(sub, loose, tight)):
        ref = s.solve_batch(Y, wa.LsmrOptions(tol=1e-12))
        assert all(rt.converged), p
        more += sum(int(a > b) for a, b in zip(rt.iterations, rl.iterations))
        for j in range(Y.shape[1]):
            if not shape[3]:
                assert rt.iterations[j] == ref.iterations[j], (p, j)
                assert rel(rt.demeaned[:, j], ref.demeaned[:, j]) <= 1e-13
            else:
                assert abs(rt.iterations[j] - ref.iterations[j]) <= 1, (p, j)
                assert rel(rt.demeaned[:, j], ref.demeaned[:, j]) <= 1e-9
    assert more > 0


def test_many_maxiter_and_zero_rhs(gpu, problems):
    s, Y = problems[0]
    Z = np.zeros_like(Y)
    opts = wa.LsmrOptions(tol=1e-14, maxiter=5)
    wa.set_many_shape(4, 2, True, True)
    try:
        r_cap, r_zero = wa.solve_many([s, s], [Y, Z], opts)
    finally:
        wa.set_many_shape(0)
    ref = s.solve_batch(Y, opts)
    assert not any(r_cap.converged) and r_cap.iterations == ref.iterations
    assert max(rel(r_cap.demeaned[:, j], ref.demeaned[:, j]) for j in range(Y.shape[1])) <= 1e-10
    assert all(r_zero.converged) and np.all(r_zero.demeaned == 0.0)


def test_many_fresh_process_runs_the_autotuner(gpu):
    """A process that never pinned a shape lets the autotuner choose (regression: the pinned
    shape used to start out as one stream without packing)."""
    import subprocess
    import sys
    code = ("import numpy as np, within_amd as wa\n"
            "from tests._configs import cfg4_problems\n"
            "pr = cfg4_problems(seed=5, n_problems=6, lo=3.3, hi=3.8)\n"
            "r = wa.solve_many([wa.Solver(c) for c, _ in pr], [Y for _, Y in pr])\n"
            "assert all(all(x.converged) for x in r)\n"
            "print('shape', wa.last_many_shape())\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240,
                         env={**__import__("os").environ, "WITHIN_AMD_VERBOSE": "1"})
    assert out.returncode == 0, out.stderr[-2000:]
    assert "autotune" in out.stderr
