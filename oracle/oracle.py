"""CPU oracle for the fixed-effects projection -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the product (``within_amd``) never does.

Two independent checkers live here:

* ``OracleLib`` -- ctypes wrapper over ``within_oracle.c``, a plain-C restatement of
  the reference's design layout, D / D' operators, (modified) LSMR and additive
  Schwarz apply (each C function cites the reference file:line it follows).
* ``exact_demean`` / ``normal_eq_residual`` -- numpy/scipy closed forms that do not
  share any code with either the C restatement or the product: the exact
  least-squares projection z - D x (gauge free) and the bincount normal-equation
  residual ||D'W(z - Dx)|| / ||D'Wz|| (reference tests/test_solve.py:23-43).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libwithin_oracle.so")

STOP_NAMES = {
    0: "ZeroRhs",
    1: "InitialNormalEquationResidualZero",
    2: "ResidualTolerance",
    3: "NormalEquationTolerance",
    5: "MaxIterations",
}


def build() -> str:
    """Compile the C oracle (gcc, OpenMP) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _Design(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("T", C.c_uint64),
        ("codes", C.c_void_p),
        ("n_levels", C.c_void_p),
        ("offset", C.c_void_p),
        ("ncols", C.c_void_p),
        ("colload", C.c_void_p),
        ("slopes", C.c_void_p),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("stop_reason", C.c_int32),
        ("converged", C.c_int32),
        ("residual_norm", C.c_double),
        ("normal_eq_residual", C.c_double),
    ]


class _Schwarz(C.Structure):
    _fields_ = [
        ("n_dofs", C.c_uint64),
        ("n_comp", C.c_uint64),
        ("slot_start", C.c_void_p),
        ("ne", C.c_void_p),
        ("nk", C.c_void_p),
        ("glob", C.c_void_p),
        ("pw", C.c_void_p),
        ("inv_de", C.c_void_p),
        ("c_ptr", C.c_void_p),
        ("c_idx", C.c_void_p),
        ("c_val", C.c_void_p),
        ("fac_kind", C.c_void_p),
        ("fac_off", C.c_void_p),
        ("fac_val", C.c_void_p),
        ("fac_ptr", C.c_void_p),
        ("fac_idx", C.c_void_p),
        ("piv_off", C.c_void_p),
        ("piv", C.c_void_p),
        ("fold", C.c_void_p),
        ("gauge", C.c_void_p),
        ("cover_off", C.c_void_p),
        ("cover_perm", C.c_void_p),
    ]


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


@dataclass
class OracleDesign:
    """A design in the oracle's layout (internal row order = caller order here;
    the locality sort is applied explicitly with ``permuted``)."""

    codes: np.ndarray  # (n, T) uint32, stored F-order
    ncols: np.ndarray  # (T,) int32
    colload: np.ndarray  # (sum ncols,) int64: -1 intercept, else slope column
    slopes: Optional[np.ndarray]  # (n, S) float64 F-order
    n_levels: np.ndarray
    offset: np.ndarray
    sorted: np.ndarray
    dominant: int
    n_dofs: int

    @property
    def n(self) -> int:
        return self.codes.shape[0]

    @property
    def T(self) -> int:
        return self.codes.shape[1]


class OracleLib:
    def __init__(self, path: str = _LIB_PATH):
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        L = self.lib
        L.wo_stable_argsort.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]
        L.wo_design_layout.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        L.wo_design_layout.restype = C.c_int64
        L.wo_gather.argtypes = [C.POINTER(_Design), C.c_void_p, C.c_void_p, C.c_void_p]
        L.wo_scatter.argtypes = [C.POINTER(_Design), C.c_void_p, C.c_void_p, C.c_void_p]
        L.wo_build_diagonal.argtypes = [C.POINTER(_Design), C.c_void_p, C.c_void_p]
        L.wo_build_diagonal.restype = C.c_int64
        L.wo_vec_norm.argtypes = [C.c_void_p, C.c_uint64]
        L.wo_vec_norm.restype = C.c_double
        L.wo_lsmr_dense.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_int32, C.c_void_p,
                                    C.c_double, C.c_int64, C.c_int64, C.c_void_p, C.POINTER(_Result)]
        L.wo_lsmr_dense.restype = C.c_int
        L.wo_solve_rhs.argtypes = [C.POINTER(_Design), C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                   C.POINTER(_Schwarz), C.c_double, C.c_int64, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.POINTER(_Result)]
        L.wo_solve_rhs.restype = C.c_int
        L.wo_solve_batch.argtypes = [C.POINTER(_Design), C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32,
                                     C.c_void_p, C.POINTER(_Schwarz), C.c_double, C.c_int64, C.c_int64,
                                     C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(_Result)]
        L.wo_solve_batch.restype = C.c_int
        L.wo_schwarz_apply.argtypes = [C.POINTER(_Schwarz), C.c_void_p, C.c_void_p]
        L.wo_schwarz_apply.restype = C.c_int
        L.wo_omp_max_threads.restype = C.c_int

    # -- design ---------------------------------------------------------------
    def stable_argsort(self, key: np.ndarray, n_levels: int) -> np.ndarray:
        key = np.ascontiguousarray(key, dtype=np.uint32)
        perm = np.empty(key.size, dtype=np.uint32)
        self.lib.wo_stable_argsort(_p(key), key.size, int(n_levels), _p(perm))
        return perm

    def design(self, codes: np.ndarray, ncols: Sequence[int] | None = None,
               colload: Sequence[int] | None = None, slopes: Optional[np.ndarray] = None) -> OracleDesign:
        codes = np.asfortranarray(np.asarray(codes, dtype=np.uint32).reshape(codes.shape[0], -1))
        T = codes.shape[1]
        ncols = np.ascontiguousarray(np.ones(T) if ncols is None else ncols, dtype=np.int32)
        colload = np.ascontiguousarray(-np.ones(int(ncols.sum())) if colload is None else colload,
                                       dtype=np.int64)
        if slopes is not None:
            slopes = np.asfortranarray(np.asarray(slopes, dtype=np.float64).reshape(codes.shape[0], -1))
        n_levels = np.zeros(T, np.uint64)
        offset = np.zeros(T, np.uint64)
        srt = np.zeros(T, np.uint8)
        dom = C.c_int64(-1)
        nd = self.lib.wo_design_layout(_p(codes), codes.shape[0], T, _p(ncols), _p(n_levels), _p(offset),
                                       _p(srt), C.byref(dom))
        if nd < 0:
            raise ValueError("empty observations")
        return OracleDesign(codes, ncols, colload, slopes, n_levels, offset, srt.astype(bool), int(dom.value),
                            int(nd))

    def permuted(self, d: OracleDesign, perm: np.ndarray) -> OracleDesign:
        """Row-permuted copy (ObservationFrame::permuted, observation.rs:96-110)."""
        slopes = None if d.slopes is None else d.slopes[perm]
        return self.design(d.codes[perm], d.ncols, d.colload, slopes)

    def _cdesign(self, d: OracleDesign) -> _Design:
        return _Design(d.n, d.T, _p(d.codes), _p(d.n_levels), _p(d.offset), _p(d.ncols), _p(d.colload),
                       _p(d.slopes))

    def gather(self, d: OracleDesign, x: np.ndarray, scale: Optional[np.ndarray] = None) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(d.n)
        cd = self._cdesign(d)
        self.lib.wo_gather(C.byref(cd), _p(x), _p(None if scale is None else np.ascontiguousarray(scale)),
                           _p(out))
        return out

    def scatter(self, d: OracleDesign, u: np.ndarray, sw: Optional[np.ndarray] = None) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(d.n_dofs)
        cd = self._cdesign(d)
        self.lib.wo_scatter(C.byref(cd), _p(u), _p(None if sw is None else np.ascontiguousarray(sw)), _p(out))
        return out

    def build_diagonal(self, d: OracleDesign, w: Optional[np.ndarray] = None) -> np.ndarray:
        out = np.empty(d.n_dofs)
        cd = self._cdesign(d)
        rc = self.lib.wo_build_diagonal(C.byref(cd), _p(None if w is None else np.ascontiguousarray(w)),
                                        _p(out))
        if rc < 0:
            raise ValueError(f"singular diagonal at {-1 - rc}")
        return out

    def vec_norm(self, v: np.ndarray) -> float:
        v = np.ascontiguousarray(v, dtype=np.float64)
        return self.lib.wo_vec_norm(_p(v), v.size)

    # -- LSMR on dense fixture operators ----------------------------------------
    def lsmr_dense(self, a: np.ndarray, b: np.ndarray, tol: float = 1e-8, maxiter: int = 1000,
                   prec: str = "none", prec_data: Optional[np.ndarray] = None, local_size: int = 0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        m, n = a.shape
        kind = {"none": 0, "diag": 1, "dense": 2, "identity": 3}[prec]
        pd = None if prec_data is None else np.ascontiguousarray(prec_data, dtype=np.float64)
        x = np.empty(n)
        res = _Result()
        rc = self.lib.wo_lsmr_dense(_p(a), m, n, _p(b), kind, _p(pd), tol, maxiter, local_size, _p(x),
                                    C.byref(res))
        return rc, x, res

    # -- full solves -----------------------------------------------------------
    def solve_batch(self, d: OracleDesign, Y: np.ndarray, w: Optional[np.ndarray] = None, prec: str = "off",
                    inv_diag: Optional[np.ndarray] = None, schwarz: Optional["OracleSchwarz"] = None,
                    tol: float = 1e-8, maxiter: int = 1000, local_size: int = 0, threads: int = 0):
        """Solve in the design's row order (caller permutes); Y is (n, k)."""
        Y = np.asfortranarray(np.asarray(Y, dtype=np.float64).reshape(d.n, -1))
        k = Y.shape[1]
        sw = None if w is None else np.ascontiguousarray(np.sqrt(w))
        kind = {"off": 0, "diagonal": 1, "schwarz": 4}[prec]
        if kind == 1 and inv_diag is None:
            inv_diag = self.build_diagonal(d, w)
        X = np.empty((d.n_dofs, k), order="F")
        Dm = np.empty((d.n, k), order="F")
        res = (_Result * max(k, 1))()
        cd = self._cdesign(d)
        sch = schwarz.struct if schwarz is not None else None
        rc = self.lib.wo_solve_batch(C.byref(cd), _p(sw), _p(Y), k, kind, _p(inv_diag),
                                     C.byref(sch) if sch is not None else None, tol, maxiter, local_size,
                                     threads, _p(X), _p(Dm), res)
        if rc != 0:
            raise RuntimeError(f"oracle solve failed rc={rc}")
        return X, Dm, [res[j] for j in range(k)]


class OracleSchwarz:
    """Holds numpy arrays backing a ``wo_schwarz`` struct (exported plain components)."""

    def __init__(self, lib: OracleLib, n_dofs: int, arrays: dict):
        self.lib = lib
        self.arrays = {}
        dtypes = dict(slot_start=np.uint64, ne=np.uint64, nk=np.uint64, glob=np.uint32, pw=np.float64,
                      inv_de=np.float64, c_ptr=np.uint64, c_idx=np.uint32, c_val=np.float64,
                      fac_kind=np.int32, fac_off=np.uint64, fac_val=np.float64, fac_ptr=np.uint64,
                      fac_idx=np.uint32, piv_off=np.uint64, piv=np.float64, fold=np.float64, gauge=np.int32,
                      cover_off=np.uint64, cover_perm=np.uint32)
        if arrays.get("fold") is None:
            raise ValueError("OracleSchwarz needs the per-slot coordinate factors `fold`")
        if arrays.get("gauge") is None:
            raise ValueError("OracleSchwarz needs the per-component `gauge`")
        for key, dt in dtypes.items():
            a = arrays.get(key)
            if a is None or len(a) == 0:
                a = np.zeros(1, dt)
            self.arrays[key] = np.ascontiguousarray(a, dtype=dt)
        A = self.arrays
        self.struct = _Schwarz(n_dofs, len(A["ne"]) if arrays.get("ne") is not None else 0,
                               *[_p(A[k]) for k in dtypes])

    def apply(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.struct.n_dofs)
        self.lib.lib.wo_schwarz_apply(C.byref(self.struct), _p(r), _p(z))
        return z


# ---------------------------------------------------------------------------
# Closed-form checkers (numpy / scipy), independent of every C implementation
# ---------------------------------------------------------------------------

def sparse_design(codes: np.ndarray, effects: Optional[list] = None):
    """scipy.sparse D in the term-major layout (offset + c*L + level).

    ``effects``: optional list of (intercept: bool, slopes: list[np.ndarray]) per term.
    """
    import scipy.sparse as sp

    codes = np.asarray(codes).reshape(np.asarray(codes).shape[0], -1)
    n, T = codes.shape
    blocks = []
    for t in range(T):
        lev = codes[:, t].astype(np.int64)
        L = int(lev.max()) + 1
        cols = []
        if effects is None:
            cols = [np.ones(n)]
        else:
            icpt, slopes = effects[t]
            if icpt:
                cols.append(np.ones(n))
            cols.extend(np.asarray(s, dtype=np.float64) for s in slopes)
        for load in cols:
            blocks.append(sp.csr_matrix((load, (np.arange(n), lev)), shape=(n, L)))
    return sp.hstack(blocks).tocsr()


def exact_demean(D, Y: np.ndarray, w: Optional[np.ndarray] = None, dense_max: int = 6000) -> np.ndarray:
    """Exact weighted projection residual z - D x with x = argmin ||W^{1/2}(z - Dx)||.

    Gauge free: the demeaned output is unique even though x is not.  Dense normal
    equations with a pseudo-inverse (eigendecomposition) below ``dense_max`` DOFs,
    else scipy's LSMR driven to machine precision on column-scaled D.
    """
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    Y = np.asarray(Y, dtype=np.float64)
    one = Y.ndim == 1
    Y2 = Y.reshape(Y.shape[0], -1)
    sw = np.ones(D.shape[0]) if w is None else np.sqrt(np.asarray(w, dtype=np.float64))
    A = sp.diags(sw) @ D
    out = np.empty_like(Y2)
    if D.shape[1] <= dense_max:
        G = (A.T @ A).toarray()
        evals, evecs = np.linalg.eigh(G)
        keep = evals > evals.max() * 1e-13
        for j in range(Y2.shape[1]):
            rhs = A.T @ (sw * Y2[:, j])
            coef = evecs[:, keep].T @ rhs / evals[keep]
            x = evecs[:, keep] @ coef
            # one step of iterative refinement in the normal equations
            r = sw * Y2[:, j] - A @ x
            g = A.T @ r
            x = x + evecs[:, keep] @ ((evecs[:, keep].T @ g) / evals[keep])
            out[:, j] = Y2[:, j] - D @ x
    else:
        colnorm = np.sqrt(np.asarray(A.multiply(A).sum(axis=0))).ravel()
        colnorm[colnorm == 0] = 1.0
        As = A @ sp.diags(1.0 / colnorm)
        for j in range(Y2.shape[1]):
            b = sw * Y2[:, j]
            x = spl.lsmr(As, b, atol=1e-15, btol=1e-15, maxiter=20000)[0]
            r = b - As @ x
            x2 = spl.lsmr(As, r, atol=1e-15, btol=1e-15, maxiter=20000)[0]
            x = (x + x2) / colnorm
            out[:, j] = Y2[:, j] - D @ x
    return out[:, 0] if one else out


def normal_eq_residual(codes_or_D, z: np.ndarray, demeaned: np.ndarray, w: Optional[np.ndarray] = None):
    """||D'W(z - Dx)|| / ||D'W z|| per column (the demeaned output is z - Dx).

    With an intercept-only design this is the reference's bincount oracle
    (tests/test_solve.py:23-43); a sparse D handles slopes.
    """
    z2 = np.asarray(z, dtype=np.float64).reshape(np.asarray(z).shape[0], -1)
    d2 = np.asarray(demeaned, dtype=np.float64).reshape(z2.shape)
    wv = np.ones(z2.shape[0]) if w is None else np.asarray(w, dtype=np.float64)
    out = []
    for j in range(z2.shape[1]):
        if isinstance(codes_or_D, np.ndarray):
            codes = codes_or_D.reshape(codes_or_D.shape[0], -1)
            num2 = 0.0
            den2 = 0.0
            for t in range(codes.shape[1]):
                lev = codes[:, t].astype(np.int64)
                num2 += float(np.sum(np.bincount(lev, weights=wv * d2[:, j]) ** 2))
                den2 += float(np.sum(np.bincount(lev, weights=wv * z2[:, j]) ** 2))
        else:
            D = codes_or_D
            num2 = float(np.sum((D.T @ (wv * d2[:, j])) ** 2))
            den2 = float(np.sum((D.T @ (wv * z2[:, j])) ** 2))
        out.append(np.sqrt(num2) / np.sqrt(den2) if den2 > 0 else 0.0)
    return np.array(out)
