"""Test helpers: decode a serialized preconditioner into the oracle's Schwarz layout."""
import struct

import numpy as np

from oracle.oracle import OracleSchwarz

_FIELDS = [("slot_start", np.uint64), ("ne", np.uint32), ("nk", np.uint32), ("fac_kind", np.uint8),
           ("gauge", np.uint8), ("glob", np.uint32), ("pw", np.float64), ("fold", np.float64),
           ("inv_de", np.float64), ("c_ptr", np.uint64),
           ("c_idx", np.uint32), ("c_val", np.float64), ("dense_off", np.uint64), ("dense", np.float64),
           ("l_row_off", np.uint64), ("l_ptr", np.uint64), ("l_idx", np.uint32), ("l_val", np.float64),
           ("piv_off", np.uint64), ("piv", np.float64), ("cover_off", np.uint64), ("cover_perm", np.uint32)]


def whiten(levels, z, w=None):
    """One slope per level centred and normalised (weighted), as the solver reparametrises a
    single-slope effect with an intercept (crates/within/src/solver/reparam.rs)."""
    w = np.ones(len(z)) if w is None else w
    L = int(levels.max()) + 1
    ws = np.bincount(levels, w, L)
    mu = np.bincount(levels, w * z, L) / np.where(ws > 0, ws, 1.0)
    c = z - mu[levels]
    nrm = np.sqrt(np.bincount(levels, w * c * c, L))
    return np.where(nrm[levels] > 0, c / np.where(nrm > 0, nrm, 1.0)[levels], 0.0)


def parse_precond(b: bytes) -> dict:
    """Decode the library's preconditioner wire format (magic WXPC, version 2)."""
    assert b[:4] == b"WXPC"
    ver, kind, n = struct.unpack_from("<IIQ", b, 4)
    pos = 20
    out = {"version": ver, "kind": kind, "n": n}
    if kind == 0:
        m, = struct.unpack_from("<Q", b, pos)
        out["inv_diag"] = np.frombuffer(b, np.float64, m, pos + 8)
        return out
    for name, dt in _FIELDS:
        m, = struct.unpack_from("<Q", b, pos)
        pos += 8
        arr = np.frombuffer(b, dt, m, pos)
        pos += m * np.dtype(dt).itemsize
        out[name] = arr
    assert pos == len(b)
    return out


def oracle_schwarz_from_bytes(lib, b: bytes) -> OracleSchwarz:
    """The exported Schwarz data in the oracle's layout: dense kept-block operators become
    oracle kind 2, sparse LDL' factors kind 1, Gremban-cover factors kind 3."""
    p = parse_precond(b)
    assert p["kind"] == 1
    kinds = np.array([2, 1, 3], np.int32)[p["fac_kind"]]
    nd = len(p["dense"])
    fac_val = np.concatenate([p["dense"], p["l_val"]])
    fac_ptr = p["l_ptr"].astype(np.uint64) + np.uint64(nd)
    fac_off = np.where(p["fac_kind"] == 0, p["dense_off"], p["l_row_off"]).astype(np.uint64)
    arrays = dict(slot_start=p["slot_start"], ne=p["ne"], nk=p["nk"], glob=p["glob"], pw=p["pw"],
                  inv_de=p["inv_de"], c_ptr=p["c_ptr"], c_idx=p["c_idx"], c_val=p["c_val"], fac_kind=kinds,
                  fac_off=fac_off, fac_val=fac_val, fac_ptr=fac_ptr,
                  fac_idx=np.concatenate([np.zeros(nd, np.uint32), p["l_idx"]]), piv_off=p["piv_off"],
                  piv=p["piv"], fold=p["fold"], gauge=p["gauge"].astype(np.int32), cover_off=p["cover_off"],
                  cover_perm=p["cover_perm"])
    return OracleSchwarz(lib, p["n"], arrays)
