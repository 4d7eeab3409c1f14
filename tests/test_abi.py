"""C-ABI and host-logic checks that need no GPU (CPU only)."""
import os
import re

import numpy as np
import pytest

import within_amd
from within_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    return sorted(set(re.findall(r"\b(wx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/*.h but not exported"


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_defaults_match_reference():
    # crates/within/src/config.rs:74-131, 170-177
    import ctypes as C
    c = _lib.PrecondConfig()
    _lib.lib().wx_precond_config_default(C.byref(c))
    assert (c.split_merge, c.schur_split, c.approx_chol_seed, c.schur_seed) == (2, 1, 0, 0)
    assert c.scaling_tolerance == 1e-9 and c.scaling_max_sweeps == 2048
    o = _lib.LsmrOpts()
    _lib.lib().wx_lsmr_options_default(C.byref(o))
    assert (o.tol, o.maxiter, o.local_size) == (1e-8, 1000, 0)
    opts = within_amd.LsmrOptions()
    assert (opts.tol, opts.maxiter, opts.local_size) == (1e-8, 1000, None)


def test_python_type_errors_before_any_device_call():
    # dtype / ndim errors are TypeError naming the dtype (crates/within-py/src/convert.rs:47-84)
    with pytest.raises(TypeError, match="uint32"):
        within_amd.solve(np.zeros((4, 2), dtype=np.int64), np.zeros(4))
    with pytest.raises(TypeError, match="float64"):
        within_amd.solve(np.zeros((4, 2), dtype=np.uint32), np.zeros(4, dtype=np.float32))
    with pytest.raises(TypeError):
        within_amd.solve("not a design", np.zeros(4))
    with pytest.raises(TypeError, match="options must be LsmrOptions"):
        within_amd.solve(np.zeros((4, 2), dtype=np.uint32), np.zeros(4), options=3)
    with pytest.raises(TypeError, match="preconditioner must be"):
        within_amd.solve(np.zeros((4, 2), dtype=np.uint32), np.zeros(4), preconditioner="fast")
    with pytest.raises(ValueError):
        within_amd.Effect(np.zeros(3, np.uint32), False)
    with pytest.raises(ValueError, match="slope 0"):
        within_amd.Effect(np.zeros(3, np.uint32), True, [np.array([1.0, np.nan, 2.0])])
    with pytest.raises(ValueError, match="split"):
        within_amd.config.ApproxSchurConfig(split=0)


def test_solver_requires_device_without_gpu():
    n, _ = _lib.device_info()
    if n > 0:
        pytest.skip("a GPU is visible; the no-device error path does not apply")
    with pytest.raises(RuntimeError, match="no AMD GPU"):
        within_amd.Solver(np.zeros((4, 2), dtype=np.uint32))


def test_preconditioner_bytes_rejected_without_device():
    # deserialisation is host-only and validates untrusted input
    for bad in [b"", b"WXPC", b"garbage bytes here", b"WXPC\x01\x00\x00\x00\x07\x00\x00\x00"]:
        with pytest.raises(ValueError, match="failed to deserialize preconditioner"):
            within_amd.Preconditioner(bad)


def test_layout_round_trip():
    lay = within_amd.CoefficientLayout(np.array([0, 6, 9]), np.array([2, 3, 2]), np.array([3, 1, 1]), 11)
    for i in range(11):
        t, l, c = lay.address(i)
        assert lay.index(t, l, c) == i
    with pytest.raises(IndexError):
        lay.index(0, 2, 0)
    with pytest.raises(IndexError):
        lay.address(11)
    with pytest.raises(IndexError):
        lay.n_levels(3)
