"""Python surface of the MI355X fixed-effects projection.

Mirrors the reference's PyO3 module ``within._within`` (crates/within-py/src/lib.rs:26-47,
api.rs, config.rs, results.rs, convert.rs): the same class and function names, keyword
names, defaults, error types and output layouts.  Every solve runs on the GPU through
``libwithin_amd.so`` (ctypes releases the GIL for the duration of each native call).

Additive fields beyond the reference: ``LsmrOptions.precision`` ("fp64" or "mixed": FP32
preconditioner coefficients, FP64 Krylov), and per-RHS ``true_residual`` (the FP64
normal-equation residual ||D'W(z - Dx)|| / ||D'Wz||; with slope covariates, D holds the
per-level reparametrised slope columns the solver works in) / ``stop_reason`` on the results.
"""
from __future__ import annotations

import ctypes as C
import struct
import sys
import time
import warnings
from typing import Optional, Sequence

import numpy as np

from . import _lib

__all__ = [
    "PreconditionerConfig", "ReductionStrategy", "LsmrOptions", "UnidentifiedDirection", "CoefficientLayout",
    "SolveResult", "BatchSolveResult", "Effect", "solve", "solve_batch", "Preconditioner", "Solver",
    "ApproxCholConfig", "ApproxSchurConfig", "Schur", "ScalingConfig", "LocalSolverConfig", "AdditiveSchwarz",
]


# ---------------------------------------------------------------------------
# configuration (crates/within-py/src/config.rs:18-311)
# ---------------------------------------------------------------------------

class PreconditionerConfig:
    """Preconditioner selection shortcut: ``Additive`` (default), ``Off``, ``Diagonal``.

    Not an Enum: the members are class attributes."""

    __slots__ = ("_kind", "_name")

    def __init__(self, kind: int, name: str):
        self._kind = kind
        self._name = name

    def __repr__(self) -> str:
        return f"PreconditionerConfig.{self._name}"

    def __eq__(self, other) -> bool:
        return isinstance(other, PreconditionerConfig) and other._kind == self._kind

    def __hash__(self) -> int:
        return hash(("PreconditionerConfig", self._kind))


PreconditionerConfig.Additive = PreconditionerConfig(3, "Additive")
PreconditionerConfig.Off = PreconditionerConfig(1, "Off")
PreconditionerConfig.Diagonal = PreconditionerConfig(2, "Diagonal")


class ReductionStrategy:
    """Combine strategy for additive Schwarz: ``Auto``, ``AtomicScatter``, ``ParallelReduction``.

    Accepted for API compatibility; the device combine is always a deterministic gather."""

    __slots__ = ("_v", "_name")

    def __init__(self, v: int, name: str):
        self._v = v
        self._name = name

    def __repr__(self) -> str:
        return f"ReductionStrategy.{self._name}"

    def __eq__(self, other) -> bool:
        return isinstance(other, ReductionStrategy) and other._v == self._v

    def __hash__(self) -> int:
        return hash(("ReductionStrategy", self._v))


ReductionStrategy.Auto = ReductionStrategy(0, "Auto")
ReductionStrategy.AtomicScatter = ReductionStrategy(1, "AtomicScatter")
ReductionStrategy.ParallelReduction = ReductionStrategy(2, "ParallelReduction")


class LsmrOptions:
    """LSMR tuning (crates/within/src/config.rs:160-178): ``tol=1e-8``, ``maxiter=1000``,
    ``local_size=None``; plus ``precision`` ("fp64" or "mixed") and ``true_residual`` (default
    True: each result's ``true_residual`` is computed, one more D' pass over the demeaned output;
    False leaves exactly the reference's work and reports it as NaN)."""

    __slots__ = ("_tol", "_maxiter", "_local_size", "_precision", "_true_residual")

    def __init__(self, tol: float = 1e-8, maxiter: int = 1000, local_size: Optional[int] = None,
                 precision: str = "fp64", true_residual: bool = True):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if int(maxiter) < 0:
            raise OverflowError("maxiter must be non-negative")
        if local_size is not None and int(local_size) < 0:
            raise OverflowError("local_size must be non-negative")
        self._tol = float(tol)
        self._maxiter = int(maxiter)
        self._local_size = None if local_size is None else int(local_size)
        self._precision = precision
        self._true_residual = bool(true_residual)

    tol = property(lambda self: self._tol)
    maxiter = property(lambda self: self._maxiter)
    local_size = property(lambda self: self._local_size)
    precision = property(lambda self: self._precision)
    true_residual = property(lambda self: self._true_residual)

    def __repr__(self) -> str:
        return (f"LsmrOptions(tol={self._tol}, maxiter={self._maxiter}, local_size={self._local_size}, "
                f"precision={self._precision!r}, true_residual={self._true_residual})")

    def _native(self) -> _lib.LsmrOpts:
        o = _lib.LsmrOpts()
        _lib.lib().wx_lsmr_options_default(C.byref(o))
        o.tol = self._tol
        o.maxiter = self._maxiter
        o.local_size = self._local_size or 0
        o.precision = 1 if self._precision == "mixed" else 0
        o.true_residual = 1 if self._true_residual else 0
        return o


class ApproxCholConfig:
    __slots__ = ("_seed", "_split_merge")

    def __init__(self, seed: int = 0, split_merge: Optional[int] = None):
        if int(seed) < 0 or (split_merge is not None and int(split_merge) < 0):
            raise OverflowError("seed and split_merge must be non-negative")
        self._seed = int(seed)
        self._split_merge = None if split_merge is None else int(split_merge)

    seed = property(lambda self: self._seed)
    split_merge = property(lambda self: self._split_merge)

    def __repr__(self) -> str:
        return f"ApproxCholConfig(seed={self._seed}, split_merge={self._split_merge})"


class ApproxSchurConfig:
    __slots__ = ("_seed", "_split")

    def __init__(self, seed: int = 0, split: int = 1):
        if int(split) == 0:
            raise ValueError("split must be >= 1")
        if int(seed) < 0 or int(split) < 0:
            raise OverflowError("seed and split must be non-negative")
        self._seed = int(seed)
        self._split = int(split)

    seed = property(lambda self: self._seed)
    split = property(lambda self: self._split)

    def __repr__(self) -> str:
        return f"ApproxSchurConfig(seed={self._seed}, split={self._split})"


class Schur:
    __slots__ = ("_exact", "_cfg")

    def __init__(self, exact: bool, cfg: Optional[ApproxSchurConfig]):
        self._exact = exact
        self._cfg = cfg

    @staticmethod
    def approximate(config: Optional[ApproxSchurConfig] = None) -> "Schur":
        return Schur(False, config or ApproxSchurConfig())

    @staticmethod
    def exact() -> "Schur":
        return Schur(True, None)

    def __repr__(self) -> str:
        if self._exact:
            return "Schur.exact()"
        return f"Schur.approximate(seed={self._cfg.seed}, split={self._cfg.split})"


class ScalingConfig:
    __slots__ = ("_tolerance", "_max_sweeps", "_on_failure")

    def __init__(self, tolerance: Optional[float] = None, max_sweeps: Optional[int] = None,
                 on_failure: Optional[str] = None):
        tol = 1e-9 if tolerance is None else float(tolerance)
        if not (np.isfinite(tol) and tol >= 0):
            raise ValueError("tolerance must be finite and non-negative")
        of = "warn" if on_failure is None else str(on_failure)
        if of not in ("warn", "error"):
            raise ValueError("on_failure must be 'warn' or 'error'")
        self._tolerance = tol
        self._max_sweeps = 2048 if max_sweeps is None else int(max_sweeps)
        self._on_failure = of

    tolerance = property(lambda self: self._tolerance)
    max_sweeps = property(lambda self: self._max_sweeps)
    on_failure = property(lambda self: self._on_failure)


class LocalSolverConfig:
    """Local solver: Schur reduction + factorisation of the kept block.

    A kept block of at most ``dense_threshold`` levels (default 24, as in the reference) gets
    the exact dense factor, approximate Cholesky otherwise.  ``dense_policy`` (an addition of
    this library) widens that rule: with ``"device"`` (the default) a larger kept block is
    also factored densely where the device cost model finds the dense apply cheaper (up to
    4,096 kept levels, or 16,384 when the reduced graph is at least 0.5 % dense);
    ``"reference"`` applies the threshold alone.  The choice depends on the config and the
    design only."""

    __slots__ = ("_approx_chol", "_schur", "_dense_threshold", "_scaling", "_dense_policy")

    def __init__(self, approx_chol: Optional[ApproxCholConfig] = None, schur: Optional[Schur] = None,
                 dense_threshold: Optional[int] = None, scaling: Optional[ScalingConfig] = None,
                 dense_policy: str = "device"):
        if dense_threshold is not None and int(dense_threshold) < 0:
            raise OverflowError("dense_threshold must be non-negative")
        if dense_policy not in ("device", "reference"):
            raise ValueError("dense_policy must be 'device' or 'reference'")
        self._approx_chol = approx_chol
        self._schur = schur
        self._dense_threshold = 24 if dense_threshold is None else int(dense_threshold)
        self._scaling = scaling
        self._dense_policy = dense_policy

    approx_chol = property(lambda self: self._approx_chol)
    schur = property(lambda self: self._schur)
    scaling = property(lambda self: self._scaling)
    dense_threshold = property(lambda self: self._dense_threshold)
    dense_policy = property(lambda self: self._dense_policy)

    def __repr__(self) -> str:
        return (f"LocalSolverConfig(approx_chol={self._approx_chol!r}, schur={self._schur!r}, "
                f"dense_threshold={self._dense_threshold}, scaling={self._scaling!r}, "
                f"dense_policy={self._dense_policy!r})")


class AdditiveSchwarz:
    __slots__ = ("_local_solver", "_reduction")

    def __init__(self, local_solver: Optional[LocalSolverConfig] = None,
                 reduction: ReductionStrategy = ReductionStrategy.Auto):
        if local_solver is not None and not isinstance(local_solver, LocalSolverConfig):
            raise TypeError("local_solver must be a LocalSolverConfig")
        if not isinstance(reduction, ReductionStrategy):
            raise TypeError("reduction must be a ReductionStrategy")
        self._local_solver = local_solver
        self._reduction = reduction

    local_solver = property(lambda self: self._local_solver)
    reduction = property(lambda self: self._reduction)


def _native_config(pc) -> Optional[_lib.PrecondConfig]:
    """PreconditionerConfig / AdditiveSchwarz -> wx_precond_config (config.rs:392-439)."""
    c = _lib.PrecondConfig()
    _lib.lib().wx_precond_config_default(C.byref(c))
    if pc is None:
        return None
    if isinstance(pc, PreconditionerConfig):
        c.kind = pc._kind
        return c
    if isinstance(pc, AdditiveSchwarz):
        c.kind = 3
        c.reduction = pc.reduction._v
        ls = pc.local_solver
        if ls is not None:
            if ls.approx_chol is not None:
                c.approx_chol_seed = ls.approx_chol.seed
                c.split_merge = ls.approx_chol.split_merge or 0
            if ls.schur is not None:
                c.schur = 1 if ls.schur._exact else 0
                if not ls.schur._exact:
                    c.schur_seed = ls.schur._cfg.seed
                    c.schur_split = ls.schur._cfg.split
            c.dense_threshold = ls.dense_threshold
            c.dense_policy = 1 if ls.dense_policy == "device" else 0
            if ls.scaling is not None:
                c.scaling_tolerance = ls.scaling.tolerance
                c.scaling_max_sweeps = ls.scaling.max_sweeps
                c.scaling_error_on_failure = 1 if ls.scaling.on_failure == "error" else 0
        return c
    raise TypeError("preconditioner must be PreconditionerConfig.Additive, PreconditionerConfig.Off, "
                    "PreconditionerConfig.Diagonal, AdditiveSchwarz(...), a pre-built Preconditioner, or None")


# ---------------------------------------------------------------------------
# results (crates/within-py/src/results.rs)
# ---------------------------------------------------------------------------

class UnidentifiedDirection:
    __slots__ = ("_t",)

    def __init__(self, term: int, level: int, column: int):
        object.__setattr__(self, "_t", (int(term), int(level), int(column)))

    term = property(lambda self: self._t[0])
    level = property(lambda self: self._t[1])
    column = property(lambda self: self._t[2])

    def __setattr__(self, k, v):
        raise AttributeError("UnidentifiedDirection is frozen")

    def __eq__(self, o) -> bool:
        return isinstance(o, UnidentifiedDirection) and o._t == self._t

    def __hash__(self) -> int:
        return hash(self._t)

    def __repr__(self) -> str:
        return f"UnidentifiedDirection(term={self.term}, level={self.level}, column={self.column})"


class CoefficientLayout:
    """(term, level, column) <-> flat x index (crates/within/src/solver.rs:125-200)."""

    def __init__(self, offsets: np.ndarray, n_levels: np.ndarray, n_columns: np.ndarray, n_dofs: int):
        self._off = [int(v) for v in offsets]
        self._lev = [int(v) for v in n_levels]
        self._col = [int(v) for v in n_columns]
        self._n = int(n_dofs)

    def n_dofs(self) -> int:
        return self._n

    def n_terms(self) -> int:
        return len(self._off)

    def _term(self, term: int) -> int:
        if not (0 <= term < len(self._off)):
            raise IndexError(f"term {term} out of range (n_terms={len(self._off)})")
        return term

    def n_levels(self, term: int) -> int:
        return self._lev[self._term(term)]

    def n_columns(self, term: int) -> int:
        return self._col[self._term(term)]

    def index(self, term: int, level: int, column: int) -> int:
        if 0 <= term < len(self._off) and 0 <= level < self._lev[term] and 0 <= column < self._col[term]:
            return self._off[term] + column * self._lev[term] + level
        raise IndexError(f"coefficient address (term={term}, level={level}, column={column}) out of range")

    def address(self, index: int) -> tuple:
        if not (0 <= index < self._n):
            raise IndexError(f"x index {index} out of range (n_dofs={self._n})")
        t = max(i for i, o in enumerate(self._off) if o <= index)
        w = index - self._off[t]
        return (t, w % self._lev[t], w // self._lev[t])

    def __repr__(self) -> str:
        return f"CoefficientLayout(n_terms={self.n_terms()}, n_dofs={self._n})"


class SolveResult:
    __slots__ = ("x", "unidentified", "layout", "demeaned", "converged", "iterations", "residual", "time_total",
                 "time_setup", "time_solve", "true_residual", "stop_reason")

    def __repr__(self) -> str:
        return (f"SolveResult(converged={self.converged}, iterations={self.iterations}, "
                f"residual={self.residual:.3e}, true_residual={self.true_residual:.3e})")


class BatchSolveResult:
    __slots__ = ("x", "unidentified", "layout", "demeaned", "converged", "iterations", "residual", "time_solve",
                 "time_setup", "time_total", "true_residual", "stop_reason")

    def __repr__(self) -> str:
        return f"BatchSolveResult(n_rhs={len(self.converged)}, converged={all(self.converged)})"


# ---------------------------------------------------------------------------
# design input (crates/within-py/src/api.rs:176-236, convert.rs:47-123)
# ---------------------------------------------------------------------------

def _dtype_name(obj) -> str:
    dt = getattr(obj, "dtype", None)
    return f", got dtype {dt}" if dt is not None else ""


def _f64_1d(name: str, obj) -> np.ndarray:
    if not isinstance(obj, np.ndarray) or obj.dtype != np.float64 or obj.ndim != 1:
        raise TypeError(f"{name} must be a 1-D float64 array{_dtype_name(obj)}")
    return obj


def _f64_2d(name: str, obj) -> np.ndarray:
    if not isinstance(obj, np.ndarray) or obj.dtype != np.float64 or obj.ndim != 2:
        raise TypeError(f"{name} must be a 2-D float64 array{_dtype_name(obj)}")
    return obj


def gram(A) -> np.ndarray:
    """A'A (m x m) of an n x m float64 matrix, formed on the GPU in one pass (1 <= m <= 64;
    deterministic).  A is a numpy array (copied to the device) or a float64 CUDA tensor (read in
    HBM; made column-major first if it is not)."""
    L = _lib.lib()
    torch = sys.modules.get("torch")
    if torch is not None and isinstance(A, torch.Tensor):
        if A.dtype != torch.float64 or A.dim() != 2:
            raise TypeError("A must be a 2-D float64 tensor")
        if not A.is_cuda:
            raise ValueError("A must be a CUDA (HIP) tensor")
        n, m = A.shape
        Ac = A.t().contiguous()  # (m, n) row-major: column-major n x m
        G = torch.empty((m, m), dtype=torch.float64, device=A.device)
        _lib.check(L.wx_gram_device(Ac.data_ptr(), n, m, n, G.data_ptr(),
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return G.cpu().numpy()
    A = np.asfortranarray(_f64_2d("A", A))
    n, m = A.shape
    G = np.empty((m, m), order="F")
    _lib.check(L.wx_gram(A.ctypes.data, n, m, n, G.ctypes.data))
    return G


def fwl(demeaned, rcond=None) -> np.ndarray:
    """Frisch-Waugh-Lovell coefficients after demeaning ``[y, X]`` together (README.md:55-66 of
    the reference): beta = argmin ||y~ - X~ beta|| with y~ = demeaned[:, 0] and X~ = demeaned[:, 1:].
    Formed from the normal equations X~'X~ beta = X~'y~: G = gram(demeaned) on the GPU, then the
    p x p least-squares solve on the host (numpy lstsq with ``rcond``; the minimum-norm solution
    when X~ is rank deficient)."""
    G = gram(demeaned)
    if G.shape[0] < 2:
        raise ValueError("demeaned needs the response and at least one regressor column")
    return np.linalg.lstsq(G[1:, 1:], G[1:, 0], rcond=rcond)[0]


class Effect:
    """One factor's effect: level codes, an optional intercept, and slope covariates
    (crates/within/src/domain/effect.rs:13-42)."""

    __slots__ = ("levels", "intercept", "slopes")

    def __init__(self, levels: np.ndarray, intercept: bool, slopes: Optional[Sequence[np.ndarray]] = None):
        if not isinstance(levels, np.ndarray) or levels.dtype != np.uint32 or levels.ndim != 1:
            raise TypeError(f"levels must be a 1-D uint32 array{_dtype_name(levels)}")
        sl = [] if slopes is None else list(slopes)
        sl = [_f64_1d("slopes", s) for s in sl]
        n = levels.shape[0]
        for i, s in enumerate(sl):
            if s.shape[0] != n:
                raise ValueError(f"slope {i} has {s.shape[0]} values, expected {n}")
            if not np.all(np.isfinite(s)):
                bad = int(np.flatnonzero(~np.isfinite(s))[0])
                raise ValueError(f"slope {i} loading at index {bad} must be finite, got {s[bad]}")
        if not intercept and not sl:
            raise ValueError("an effect must have an intercept or at least one slope")
        self.levels = np.ascontiguousarray(levels)
        self.intercept = bool(intercept)
        self.slopes = [np.ascontiguousarray(s) for s in sl]


def _make_design(design):
    L = _lib.lib()
    out = C.c_void_p()
    if isinstance(design, np.ndarray):
        if design.dtype != np.uint32 or design.ndim != 2:
            raise TypeError(f"design must be a 2-D uint32 array{_dtype_name(design)}")
        if design.shape[0] > 1 and design.strides[0] != design.itemsize:
            warnings.warn(
                "categories array is not F-contiguous (column-major). If the data is already sorted by the "
                "largest factor, np.asfortranarray(categories) gives faster solves; unsorted input is copied "
                "internally either way.", UserWarning, stacklevel=3)
        n, f = design.shape
        rs = design.strides[0] // design.itemsize
        cs = design.strides[1] // design.itemsize
        _lib.check(L.wx_design_from_codes(design.ctypes.data, n, f, rs, cs, C.byref(out)))
        return out
    if isinstance(design, (list, tuple)) and all(isinstance(e, Effect) for e in design):
        effects = list(design)
        n = effects[0].levels.shape[0] if effects else 0
        for e in effects:
            if e.levels.shape[0] != n:
                raise ValueError(f"column has {e.levels.shape[0]} observations, expected {n}")
        lv = (C.c_void_p * max(1, len(effects)))(*[e.levels.ctypes.data for e in effects])
        ic = np.array([e.intercept for e in effects], dtype=np.uint8)
        ns = np.array([len(e.slopes) for e in effects], dtype=np.uint64)
        flat = [s for e in effects for s in e.slopes]
        sp = (C.c_void_p * max(1, len(flat)))(*[s.ctypes.data for s in flat])
        _lib.check(L.wx_design_from_effects(n, len(effects), lv, ic.ctypes.data, ns.ctypes.data, sp, C.byref(out)))
        return out
    raise TypeError("design must be a 2-D uint32 array or a list of Effect")


# ---------------------------------------------------------------------------
# Preconditioner / Solver (crates/within-py/src/api.rs:238-347, config.rs:313-389)
# ---------------------------------------------------------------------------

class Preconditioner:
    """Pre-built preconditioner; picklable, reusable across solvers, applied on the GPU."""

    __slots__ = ("_h", "__weakref__")

    def __init__(self, data: bytes):
        if isinstance(data, C.c_void_p):  # internal: wrap a native handle
            self._h = data
            return
        if not isinstance(data, (bytes, bytearray, memoryview)):
            raise TypeError("Preconditioner(data) expects bytes")
        b = bytes(data)
        if b[:4] != b"WXPC":
            b = _from_reference_payload(b)
        out = C.c_void_p()
        buf = (C.c_uint8 * max(1, len(b))).from_buffer_copy(b if b else b"\0")
        st = _lib.lib().wx_precond_deserialize(buf, len(b), C.byref(out))
        if st != 0:
            msg = _lib.lib().wx_last_error().decode()
            raise ValueError(f"failed to deserialize preconditioner: {msg}")
        self._h = out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.lib().wx_precond_release(h)
            except Exception:
                pass

    @property
    def nrows(self) -> int:
        return int(_lib.lib().wx_precond_n(self._h))

    @property
    def ncols(self) -> int:
        return int(_lib.lib().wx_precond_n(self._h))

    def apply(self, x: np.ndarray) -> np.ndarray:
        """y = M^{-1} x (on the GPU)."""
        x = _f64_1d("x", x)
        if not x.flags.c_contiguous:
            raise ValueError("x must be contiguous")
        if x.shape[0] != self.ncols:
            raise ValueError(f"x has length {x.shape[0]} but preconditioner expects {self.ncols}")
        y = np.empty_like(x)
        _lib.check(_lib.lib().wx_precond_apply(self._h, x.ctypes.data, 1, y.ctypes.data))
        return y

    def to_bytes(self) -> bytes:
        L = _lib.lib()
        p = C.POINTER(C.c_uint8)()
        n = C.c_uint64()
        _lib.check(L.wx_precond_serialize(self._h, C.byref(p), C.byref(n)))
        try:
            return C.string_at(p, n.value)
        finally:
            L.wx_buffer_free(p)

    def __reduce__(self):
        return (Preconditioner, (self.to_bytes(),))

    def __repr__(self) -> str:
        return f"Preconditioner({_lib.lib().wx_precond_variant(self._h).decode()}, n={self.nrows})"


def _from_reference_payload(b: bytes) -> bytes:
    """Bytes pickled by the reference (postcard wire v13): a Diagonal preconditioner converts to
    this library's format; an additive Schwarz one is decoded and reported, since its local
    factors are in approx-chol's own format.  Other bytes pass through unchanged (the native
    reader reports them)."""
    from . import wire_v13

    try:
        ref = wire_v13.decode(b)
    except wire_v13.WireError as e:
        if b[:1] in (b"\x00", b"\x01") and len(b) > 1:
            raise ValueError(f"failed to deserialize preconditioner: not a valid v13 payload ({e})") from None
        return b
    if ref.variant == "Diagonal":
        inv = np.ascontiguousarray(ref.inv_diag, "<f8")
        return b"WXPC" + struct.pack("<IIQQ", 2, 0, len(inv), len(inv)) + inv.tobytes()
    raise ValueError(
        f"a reference additive Schwarz preconditioner (wire v13: {len(ref.subdomains)} subdomains, "
        f"n_dofs={ref.n_dofs}) stores its local factors in approx-chol's format, which this library "
        "does not read; rebuild it with Solver(design, preconditioner=PreconditionerConfig.Additive) "
        "or build_preconditioner(design)")


def build_preconditioner(design, weights: Optional[np.ndarray] = None, preconditioner=None) -> Optional[Preconditioner]:
    """Builds a preconditioner on the host only (no GPU needed), e.g. for offline
    construction and pickling; ``None`` when the config resolves to no preconditioner."""
    L = _lib.lib()
    if weights is not None:
        weights = np.ascontiguousarray(_f64_1d("weights", weights))
    cfg = _native_config(preconditioner)
    d = _make_design(design)
    try:
        out = C.c_void_p()
        _lib.check(L.wx_precond_build(d, None if weights is None else weights.ctypes.data,
                                      None if cfg is None else C.byref(cfg), C.byref(out)))
    finally:
        L.wx_design_release(d)
    return Preconditioner(out) if out.value else None


def _check_options(options) -> LsmrOptions:
    if options is None:
        return LsmrOptions()
    if not isinstance(options, LsmrOptions):
        raise TypeError("options must be LsmrOptions")
    return options


class Solver:
    """Persistent solver: the preconditioner is built once (host), the design uploaded
    once (HBM); each solve runs the batched LSMR on the GPU."""

    def __init__(self, design, weights: Optional[np.ndarray] = None, preconditioner=None):
        t0 = time.perf_counter()
        L = _lib.lib()
        if weights is not None:
            weights = np.ascontiguousarray(_f64_1d("weights", weights))
        prebuilt = preconditioner if isinstance(preconditioner, Preconditioner) else None
        cfg = None if prebuilt is not None else _native_config(preconditioner)
        d = _make_design(design)
        try:
            if weights is not None:
                # the weight count check precedes native validation (domain.rs:261-267)
                n_obs = (design.shape[0] if isinstance(design, np.ndarray)
                         else (design[0].levels.shape[0] if design else 0))
                if weights.shape[0] != n_obs:
                    raise ValueError(f"weights has length {weights.shape[0]}, expected {n_obs}")
            out = C.c_void_p()
            _lib.check(L.wx_solver_new(d, None if weights is None else weights.ctypes.data,
                                       None if cfg is None else C.byref(cfg),
                                       None if prebuilt is None else prebuilt._h, C.byref(out)))
        finally:
            L.wx_design_release(d)
        self._adopt(out, prebuilt is not None, t0)

    def _adopt(self, out, prebuilt: bool, t0: float):
        L = _lib.lib()
        self._h = out
        self._prebuilt = prebuilt
        self._time_setup = time.perf_counter() - t0
        nt = int(L.wx_solver_n_terms(out))
        off = np.zeros(nt, np.uint64)
        lev = np.zeros(nt, np.uint64)
        col = np.zeros(nt, np.uint64)
        L.wx_solver_layout(out, off.ctypes.data, lev.ctypes.data, col.ctypes.data)
        self._layout = CoefficientLayout(off, lev, col, int(L.wx_solver_n_dofs(out)))
        cnt = C.c_uint64(0)
        _lib.check(L.wx_solver_unidentified(out, None, C.byref(cnt)))
        tr = np.zeros(max(1, cnt.value) * 3, np.uint64)
        _lib.check(L.wx_solver_unidentified(out, tr.ctypes.data, C.byref(cnt)))
        self._unidentified = [UnidentifiedDirection(*tr[3 * i:3 * i + 3]) for i in range(cnt.value)]
        self._warnings = [L.wx_solver_warning(out, i).decode() for i in range(int(L.wx_solver_n_warnings(out)))]
        for w in self._warnings:
            warnings.warn(w, UserWarning, stacklevel=3)

    def reweight(self, weights: Optional[np.ndarray], preconditioner=None) -> "Solver":
        """A Solver for this design with new observation weights (None: unweighted), equal to
        ``Solver(design, weights, preconditioner)``, that reuses this Solver's device design
        (copied on the device) instead of rebuilding it from the codes.  ``preconditioner`` is
        a config (built anew for the new weights) or a prebuilt Preconditioner, e.g. this
        Solver's own, which stays an SPD preconditioner of the reweighted normal equations.
        Designs with slope columns are rejected (their whitened columns depend on the weights)."""
        t0 = time.perf_counter()
        L = _lib.lib()
        if weights is not None:
            weights = np.ascontiguousarray(_f64_1d("weights", weights))
            if weights.shape[0] != self.n_obs:
                raise ValueError(f"weights has length {weights.shape[0]}, expected {self.n_obs}")
        prebuilt = preconditioner if isinstance(preconditioner, Preconditioner) else None
        cfg = None if prebuilt is not None else _native_config(preconditioner)
        out = C.c_void_p()
        _lib.check(L.wx_solver_reweight(self._h, None if weights is None else weights.ctypes.data,
                                        None if cfg is None else C.byref(cfg),
                                        None if prebuilt is None else prebuilt._h, C.byref(out)))
        r = Solver.__new__(Solver)
        r._adopt(out, prebuilt is not None, t0)
        return r

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.lib().wx_solver_release(h)
            except Exception:
                pass

    @property
    def n_dofs(self) -> int:
        return int(_lib.lib().wx_solver_n_dofs(self._h))

    @property
    def n_obs(self) -> int:
        return int(_lib.lib().wx_solver_n_obs(self._h))

    @property
    def preconditioner(self) -> Optional[Preconditioner]:
        out = C.c_void_p()
        _lib.check(_lib.lib().wx_solver_preconditioner(self._h, C.byref(out)))
        return Preconditioner(out) if out.value else None

    def _run(self, Y: np.ndarray, options: LsmrOptions):
        n, k = Y.shape
        N = self.n_dofs
        X = np.empty((N, k), order="F")
        D = np.empty((n, k), order="F")
        stats = (_lib.RhsStats * max(1, k))()
        if k:
            _lib.check(_lib.lib().wx_solver_solve_batch(self._h, Y.ctypes.data, n, k, C.byref(options._native()),
                                                        X.ctypes.data, D.ctypes.data, stats))
        return X, D, [stats[j] for j in range(k)]

    def solve(self, y: np.ndarray, options: Optional[LsmrOptions] = None) -> SolveResult:
        t0 = time.perf_counter()
        options = _check_options(options)
        y = _f64_1d("y", y)
        if y.shape[0] != self.n_obs:
            raise ValueError(f"Solver::solve: response vector length ({y.shape[0]}) does not match number of "
                             f"observations ({self.n_obs})")
        X, D, st = self._run(np.asfortranarray(y.reshape(-1, 1)), options)
        r = SolveResult()
        r.x = X[:, 0].copy()
        r.demeaned = D[:, 0].copy()
        r.unidentified = list(self._unidentified)
        r.layout = self._layout
        s = st[0]
        r.converged = bool(s.converged)
        r.iterations = int(s.iterations)
        r.residual = float(s.residual)
        r.true_residual = float(s.true_residual)
        r.stop_reason = _lib.STOP_REASONS.get(int(s.stop_reason), str(s.stop_reason))
        r.time_solve = float(s.time_solve)
        r.time_total = time.perf_counter() - t0
        r.time_setup = max(0.0, r.time_total - r.time_solve)
        return r

    def solve_batch(self, Y: np.ndarray, options: Optional[LsmrOptions] = None) -> BatchSolveResult:
        t0 = time.perf_counter()
        options = _check_options(options)
        Y = _f64_2d("Y", Y)
        if Y.shape[0] != self.n_obs:
            raise ValueError(f"Y has {Y.shape[0]} rows, expected n_obs={self.n_obs}")
        X, D, st = self._run(np.asfortranarray(Y), options)
        r = BatchSolveResult()
        r.x = X
        r.demeaned = D
        r.unidentified = list(self._unidentified)
        r.layout = self._layout
        r.converged = [bool(s.converged) for s in st]
        r.iterations = [int(s.iterations) for s in st]
        r.residual = [float(s.residual) for s in st]
        r.true_residual = [float(s.true_residual) for s in st]
        r.stop_reason = [_lib.STOP_REASONS.get(int(s.stop_reason), str(s.stop_reason)) for s in st]
        r.time_solve = [float(s.time_solve) for s in st]
        r.time_setup = 0.0
        r.time_total = time.perf_counter() - t0
        return r


def solve(design, y: np.ndarray, weights: Optional[np.ndarray] = None, options: Optional[LsmrOptions] = None,
          preconditioner=None) -> SolveResult:
    """One-shot Solver(design, weights, preconditioner).solve(y, options) (solver.rs:611-626)."""
    t0 = time.perf_counter()
    options = _check_options(options)
    y = _f64_1d("y", y)
    s = Solver(design, weights, preconditioner)
    setup = time.perf_counter() - t0
    r = s.solve(y, options)
    r.time_setup += setup
    r.time_total = time.perf_counter() - t0
    return r


def solve_batch(design, Y: np.ndarray, weights: Optional[np.ndarray] = None, options: Optional[LsmrOptions] = None,
                preconditioner=None) -> BatchSolveResult:
    """One-shot Solver(...).solve_batch(Y, options) (solver.rs:632-646)."""
    t0 = time.perf_counter()
    options = _check_options(options)
    Y = _f64_2d("Y", Y)
    s = Solver(design, weights, preconditioner)
    setup = time.perf_counter() - t0
    r = s.solve_batch(Y, options)
    r.time_setup += setup
    r.time_total = time.perf_counter() - t0
    return r


def solve_many(solvers, Ys, options: Optional[LsmrOptions] = None) -> list:
    """Solves many independent problems at once: ``solvers[p].solve_batch(Ys[p], options)``
    for every p, with the problems iterating concurrently on the GPU (several streams,
    captured iteration graphs, per-problem convergence; the launch shape is chosen by an
    autotuner on the first call for a problem-count class).  Returns the list of
    BatchSolveResult.  No counterpart in the reference, whose solves run one Solver at a time."""
    t0 = time.perf_counter()
    options = _check_options(options)
    solvers = list(solvers)
    Ys = [np.asfortranarray(_f64_2d("Y", Y)) for Y in Ys]
    if len(solvers) != len(Ys):
        raise ValueError(f"{len(solvers)} solvers but {len(Ys)} right-hand-side blocks")
    P = len(solvers)
    for p, (s, Y) in enumerate(zip(solvers, Ys)):
        if not isinstance(s, Solver):
            raise TypeError("solve_many takes Solver objects")
        if Y.shape[0] != s.n_obs:
            raise ValueError(f"problem {p}: Y has {Y.shape[0]} rows, expected n_obs={s.n_obs}")
    Xs = [np.empty((s.n_dofs, Y.shape[1]), order="F") for s, Y in zip(solvers, Ys)]
    Ds = [np.empty(Y.shape, order="F") for Y in Ys]
    stats = [(_lib.RhsStats * max(1, Y.shape[1]))() for Y in Ys]
    P_ = max(P, 1)
    hs = (C.c_void_p * P_)(*[s._h for s in solvers])
    yp = (C.c_void_p * P_)(*[Y.ctypes.data for Y in Ys])
    ld = (C.c_uint64 * P_)(*[Y.shape[0] for Y in Ys])
    kk = (C.c_uint64 * P_)(*[Y.shape[1] for Y in Ys])
    xp = (C.c_void_p * P_)(*[X.ctypes.data for X in Xs])
    dp = (C.c_void_p * P_)(*[D.ctypes.data for D in Ds])
    sp = (C.c_void_p * P_)(*[C.cast(st, C.c_void_p) for st in stats])
    _lib.check(_lib.lib().wx_solve_many(hs, P, yp, ld, kk, C.byref(options._native()), xp, dp, sp))
    out = []
    total = time.perf_counter() - t0
    for s, X, D, st, Y in zip(solvers, Xs, Ds, stats, Ys):
        sl = [st[j] for j in range(Y.shape[1])]
        r = BatchSolveResult()
        r.x = X
        r.demeaned = D
        r.unidentified = list(s._unidentified)
        r.layout = s._layout
        r.converged = [bool(q.converged) for q in sl]
        r.iterations = [int(q.iterations) for q in sl]
        r.residual = [float(q.residual) for q in sl]
        r.true_residual = [float(q.true_residual) for q in sl]
        r.stop_reason = [_lib.STOP_REASONS.get(int(q.stop_reason), str(q.stop_reason)) for q in sl]
        r.time_solve = [float(q.time_solve) for q in sl]
        r.time_setup = 0.0
        r.time_total = total
        out.append(r)
    return out


def last_many_shape() -> dict:
    """The launch shape the autotuner picked for the last solve_many call."""
    s, g, gr, pk = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    _lib.lib().wx_last_many_shape(C.byref(s), C.byref(g), C.byref(gr), C.byref(pk))
    return {"streams": s.value, "iterations_per_graph": g.value, "graph": bool(gr.value), "pack": bool(pk.value),
            "tile": int(_lib.lib().wx_last_many_tile())}


def set_many_shape(streams: int = 0, iterations_per_graph: int = 1, graph: bool = True, pack: bool = False,
                   tile: int = 0) -> None:
    """Fixes the launch shape of later solve_many calls (pack: problems of one shape run as one
    block-diagonal batch; tile: rows per D' and gather work item of the packs, 0 = 64 when the
    shape is fixed); streams=0 returns the choice to the autotuner."""
    _lib.lib().wx_set_many_shape(int(streams), int(iterations_per_graph), 1 if graph else 0, 1 if pack else 0)
    _lib.lib().wx_set_many_tile(int(tile) if streams else 0)
