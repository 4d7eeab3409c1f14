"""Config 4 at its workload: 64 independent problems with n log-uniform in [1e4, 1e6], 2 or 3
factors (n/20, n/200 and 30 uniform levels), 1-8 right-hand sides each, solved by ONE solve_many
call with the autotuned launch shape, in mixed and in FP64 precision.

Every problem is held to references that do not come from the library's own solves:
  - the normal-equation residual ||D'(z - Dx)|| / ||D'z|| recomputed with numpy bincounts
    (tests/test_solve.py:23-43 of the reference) must equal the reported true residual;
  - the first column's demeaned output must match the CPU oracle (the C restatement of the
    reference's mLSMR + additive Schwarz) solving the same column at the same tol with the same
    preconditioner bytes, to 1e-8 relative (the north star's bar against the reference CPU solve);
  - problems with at most 1,500 coefficients are also held to the exact projection."""
import numpy as np
import pytest

import within_amd as wa
from oracle.oracle import exact_demean, normal_eq_residual, sparse_design
from tests._configs import cfg4_problems
from tests._helpers import oracle_schwarz_from_bytes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def workload():
    probs = cfg4_problems(seed=4, n_problems=64)
    print(f"[cfg4] {len(probs)} problems, {sum(c.shape[0] for c, _ in probs):,} rows, "
          f"{sum(Y.shape[1] for _, Y in probs)} right-hand sides", flush=True)
    solvers = [wa.Solver(c) for c, _ in probs]
    return probs, solvers


@pytest.fixture(scope="module")
def oracle_first_columns(workload, oracle):
    probs, solvers = workload
    out = []
    for (codes, Y), s in zip(probs, solvers):
        sch = oracle_schwarz_from_bytes(oracle, s.preconditioner.to_bytes())
        _, dm, res = oracle.solve_batch(oracle.design(codes), np.asfortranarray(Y[:, :1]), None, prec="schwarz",
                                        schwarz=sch, tol=1e-8)
        assert res[0].converged
        out.append(dm[:, 0])
    return out


@pytest.fixture(scope="module")
def exact(workload):
    probs, solvers = workload
    return {p: exact_demean(sparse_design(c), Y) for p, ((c, Y), s) in enumerate(zip(probs, solvers))
            if s.n_dofs <= 1500}


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_config4_full_workload_in_one_call(gpu, workload, oracle_first_columns, exact, precision):
    probs, solvers = workload
    wa.set_many_shape(0)  # the autotuner chooses
    opts = wa.LsmrOptions(tol=1e-8, precision=precision)
    res = wa.solve_many(solvers, [Y for _, Y in probs], opts)
    print(f"[cfg4] {precision}: shape {wa.last_many_shape()}", flush=True)
    worst_or = worst_ex = worst_tr = 0.0
    for p, ((codes, Y), s, r, dm_o) in enumerate(zip(probs, solvers, res, oracle_first_columns)):
        assert r.demeaned.shape == Y.shape and r.x.shape == (s.n_dofs, Y.shape[1])
        assert all(r.converged), p
        ne = np.array(normal_eq_residual(codes, Y, r.demeaned))
        tr = np.array(r.true_residual)
        assert np.all(np.abs(tr - ne) <= 1e-6 * ne + 1e-15), (p, tr, ne)
        worst_tr = max(worst_tr, float(tr.max()))
        d_or = rel(r.demeaned[:, 0], dm_o)
        worst_or = max(worst_or, d_or)
        assert d_or <= 1e-8, (p, d_or)
        if p in exact:
            for j in range(Y.shape[1]):
                d_ex = rel(r.demeaned[:, j], exact[p][:, j])
                worst_ex = max(worst_ex, d_ex)
                assert d_ex <= 1e-7, (p, j, d_ex)
    print(f"[cfg4] {precision}: worst vs oracle {worst_or:.2e}, vs exact ({len(exact)} problems) "
          f"{worst_ex:.2e}, max true residual {worst_tr:.2e}", flush=True)
