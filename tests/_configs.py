"""Seeded generators of the hot path's workloads (full sizes; tests may scale them down).

config 2: weighted solve_batch, n = 5M, A 200,000 Zipf(1.1) levels, B 10,000 and C 50
          uniform levels, weights U(0.5, 2), k = 16.
config 3: worker-firm panel, 400,000 workers x 5 periods, initial firm Zipf(1.2) over 50,000
          firms, moves w.p. 0.05 per period, year effect, firm intercept + one slope, k = 4.
config 4: 64 independent problems, n log-uniform in [1e4, 1e6], 2 or 3 factors with
          n/20, n/200 and 30 uniform levels, 1-8 right-hand sides each.
config 5: n = 1M, A 50,000 Zipf(1.1) levels, B 2,000 uniform, k = 8, w = exp(N(0, 0.5)).
"""
import numpy as np


def cfg2_data(n=5_000_000, k=16, seed=2, la=200_000, lb=10_000, lc=50):
    rng = np.random.default_rng(seed)
    p = (np.arange(la) + 1.0) ** -1.1
    p /= p.sum()
    A = rng.choice(la, n, p=p)
    B = rng.integers(0, lb, n)
    Cc = rng.integers(0, lc, n)
    codes = np.asfortranarray(np.column_stack([A, B, Cc]).astype(np.uint32))
    w = rng.uniform(0.5, 2.0, n)
    Y = np.asfortranarray(rng.standard_normal((n, k)))
    return codes, w, Y


def cfg3_data(seed=3, nw=400_000, T=5, nf=50_000, k=4):
    rng = np.random.default_rng(seed)
    p = (np.arange(nf) + 1.0) ** -1.2
    p /= p.sum()
    firm = np.empty((nw, T), np.int64)
    firm[:, 0] = rng.choice(nf, nw, p=p)
    for t in range(1, T):
        move = rng.uniform(size=nw) < 0.05
        firm[:, t] = np.where(move, rng.integers(0, nf, nw), firm[:, t - 1])
    worker = np.repeat(np.arange(nw), T).astype(np.uint32)
    firm = firm.ravel().astype(np.uint32)
    year = np.tile(np.arange(T), nw).astype(np.uint32)
    z = rng.standard_normal(nw * T)
    Y = np.asfortranarray(rng.standard_normal((nw * T, k)))
    return worker, firm, year, z, Y


def cfg4_problems(seed=4, n_problems=64, lo=4.0, hi=6.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_problems):
        n = int(round(10 ** rng.uniform(lo, hi)))
        F = 2 if rng.uniform() < 0.5 else 3
        levels = [max(1, n // 20), max(1, n // 200), 30][:F]
        codes = np.asfortranarray(np.column_stack([rng.integers(0, L, n) for L in levels]).astype(np.uint32))
        k = int(rng.integers(1, 9))
        Y = np.asfortranarray(rng.standard_normal((n, k)))
        out.append((codes, Y))
    return out


def cfg5_data(seed=5, n=1_000_000, la=50_000, lb=2_000, k=8):
    rng = np.random.default_rng(seed)
    p = (np.arange(la) + 1.0) ** -1.1
    p /= p.sum()
    A = rng.choice(la, n, p=p)
    B = rng.integers(0, lb, n)
    codes = np.asfortranarray(np.column_stack([A, B]).astype(np.uint32))
    w = np.exp(rng.normal(0.0, 0.5, n))
    Y = np.asfortranarray(rng.standard_normal((n, k)))
    mask = rng.standard_normal((n, k))
    return codes, w, Y, mask
