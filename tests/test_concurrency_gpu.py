"""One Solver shared by several Python threads (the reference's tests/test_free_threaded.py:32-58).
The bindings release the GIL for every native call, so the calls below overlap on the device; each
must give bitwise what the same call gives alone."""
import threading

import numpy as np
import pytest

import within_amd as wa
from tests._configs import cfg4_problems

pytestmark = pytest.mark.gpu


def _run_threads(fns):
    out = [None] * len(fns)
    err = []

    def work(i):
        try:
            out[i] = fns[i]()
        except BaseException as e:  # noqa: BLE001 -- reported below
            err.append(e)

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(fns))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if err:
        raise err[0]
    return out


def test_shared_solver_solve_batch_from_threads(gpu):
    rng = np.random.default_rng(11)
    n = 200_000
    codes = np.asfortranarray(np.column_stack([rng.integers(0, 10_000, n), rng.integers(0, 500, n),
                                               rng.integers(0, 12, n)]).astype(np.uint32))
    s = wa.Solver(codes, weights=rng.uniform(0.5, 2.0, n))
    Ys = [np.asfortranarray(rng.standard_normal((n, k))) for k in (1, 3, 4, 8)]
    opts = wa.LsmrOptions(tol=1e-9)
    serial = [s.solve_batch(Y, opts) for Y in Ys]
    for rep in range(3):
        print(f"[threads] solve_batch round {rep}", flush=True)
        par = _run_threads([lambda Y=Y: s.solve_batch(Y, opts) for Y in Ys])
        for a, b in zip(par, serial):
            assert a.iterations == b.iterations
            assert np.array_equal(a.demeaned, b.demeaned) and np.array_equal(a.x, b.x)


@pytest.mark.parametrize("shape", [(4, 2, True, False), (4, 2, True, True)])
def test_shared_solvers_solve_many_from_threads(gpu, shape):
    """Two threads run solve_many over the same Solvers with a graph shape pinned: each call
    leases its own streams, so no capture records the other call's launches."""
    probs = cfg4_problems(seed=43, n_problems=8, lo=3.5, hi=4.5)
    solvers = [wa.Solver(c) for c, _ in probs]
    Ys = [Y for _, Y in probs]
    Ys2 = [np.asfortranarray(Y[:, ::-1]) for Y in Ys]
    opts = wa.LsmrOptions(tol=1e-9)
    wa.set_many_shape(*shape)
    try:
        ref1 = wa.solve_many(solvers, Ys, opts)
        ref2 = wa.solve_many(solvers, Ys2, opts)
        for rep in range(3):
            print(f"[threads] solve_many {shape} round {rep}", flush=True)
            r1, r2 = _run_threads([lambda: wa.solve_many(solvers, Ys, opts),
                                   lambda: wa.solve_many(solvers, Ys2, opts)])
            for got, ref in ((r1, ref1), (r2, ref2)):
                for a, b in zip(got, ref):
                    assert a.iterations == b.iterations
                    assert np.array_equal(a.demeaned, b.demeaned)
    finally:
        wa.set_many_shape(0)
