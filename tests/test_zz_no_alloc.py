"""The allocation-free Gram entry point against the allocating one, through the C ABI.

wx_gram_device_noalloc runs the same kernels and grid as wx_gram_device with its partial sums
in a static device array, so G must agree bitwise; two calls in a row, the second smaller,
check that the array's leftover contents never leak into a result."""
import ctypes as C

import numpy as np
import pytest
import torch

from within_amd import _lib


def _gram(fn, A, m, stream):
    G = torch.empty((m, m), dtype=torch.float64, device=A.device)
    n = A.shape[1]
    _lib.check(fn(C.c_void_p(A.data_ptr()), n, m, n, C.c_void_p(G.data_ptr()), C.c_void_p(stream)))
    return G.cpu().numpy().T  # column-major m x m


@pytest.mark.gpu
def test_gram_noalloc_matches_allocating_entry(gpu):
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(77)
    for n, m in ((300_000, 24), (50_000, 7), (1_000, 32)):
        print(f"[no_alloc] n={n} m={m}", flush=True)
        A_h = rng.standard_normal((n, m))
        # column-major n x m on the device: a (m, n) row-major tensor
        A = torch.from_numpy(np.ascontiguousarray(A_h.T)).to("cuda")
        G_ref = _gram(L.wx_gram_device, A, m, stream)
        G_na = _gram(L.wx_gram_device_noalloc, A, m, stream)
        assert np.array_equal(G_na, G_ref), np.abs(G_na - G_ref).max()
        G_np = A_h.T @ A_h
        assert np.abs(G_na - G_np).max() <= 1e-12 * np.abs(G_np).max()


@pytest.mark.gpu
def test_gram_noalloc_rejects_too_many_columns(gpu):
    L = _lib.lib()
    A = torch.zeros((33, 10), dtype=torch.float64, device="cuda")
    G = torch.empty((33, 33), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="1 to 32 columns"):
        _lib.check(L.wx_gram_device_noalloc(C.c_void_p(A.data_ptr()), 10, 33, 10, C.c_void_p(G.data_ptr()), None))
