"""The Frisch-Waugh-Lovell step: gram(A) = A'A on the GPU, and fwl(demeaned) against numpy's
lstsq on the same demeaned block (the reference's README.md:55-66 example)."""
import numpy as np
import pytest
import torch

import within_amd as wa
from oracle.oracle import exact_demean, sparse_design

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m", [1, 7, 16, 64])
def test_gram_matches_numpy(gpu, m):
    rng = np.random.default_rng(100 + m)
    n = 200_003  # a ragged last chunk
    A = rng.standard_normal((n, m)) * np.exp(rng.normal(0.0, 1.0, m))
    G = wa.gram(A)
    ref = A.T @ A
    np.testing.assert_allclose(G, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.array_equal(G, G.T)


def test_gram_device_tensor_and_determinism(gpu):
    rng = np.random.default_rng(7)
    A = rng.standard_normal((50_000, 9))
    t = torch.tensor(A, device="cuda")          # row-major (n, m): made column-major inside
    G1, G2 = wa.gram(t), wa.gram(t)
    assert np.array_equal(G1, G2)
    assert np.array_equal(G1, wa.gram(np.asfortranarray(A)))
    np.testing.assert_allclose(G1, A.T @ A, rtol=1e-12, atol=1e-9)


def test_gram_edge_cases(gpu):
    assert np.array_equal(wa.gram(np.zeros((0, 3))), np.zeros((3, 3)))
    with pytest.raises(ValueError):
        wa.gram(np.zeros((10, 65)))
    with pytest.raises(ValueError):
        wa.fwl(np.zeros((10, 1)))


def test_fwl_readme_example(gpu):
    # README.md:29-45 data and 55-66 regression
    np.random.seed(1)
    n = 100_000
    fe = np.asfortranarray(np.column_stack([np.random.randint(0, 500, n).astype(np.uint32),
                                            np.random.randint(0, 200, n).astype(np.uint32)]))
    beta_true = np.array([1.0, -2.0, 0.5])
    X = np.random.randn(n, 3)
    y = X @ beta_true + np.random.randn(n)
    r = wa.solve_batch(fe, np.column_stack([y, X]), options=wa.LsmrOptions(tol=1e-10))
    beta = wa.fwl(r.demeaned)
    ref = np.linalg.lstsq(r.demeaned[:, 1:], r.demeaned[:, 0], rcond=None)[0]
    np.testing.assert_allclose(beta, ref, rtol=1e-10)
    # against the exact projection of the same data
    ex = exact_demean(sparse_design(fe), np.column_stack([y, X]))
    beta_ex = np.linalg.lstsq(ex[:, 1:], ex[:, 0], rcond=None)[0]
    np.testing.assert_allclose(beta, beta_ex, rtol=1e-8)
    assert np.abs(beta - beta_true).max() < 0.02
