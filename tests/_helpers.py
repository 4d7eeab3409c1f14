"""Test helpers: decode a serialized preconditioner into the oracle's Schwarz layout."""
import struct

import numpy as np

from oracle.oracle import OracleSchwarz

_FIELDS = [("slot_start", np.uint64), ("ne", np.uint32), ("nk", np.uint32), ("fac_kind", np.uint8),
           ("glob", np.uint32), ("pw", np.float64), ("inv_de", np.float64), ("c_ptr", np.uint64),
           ("c_idx", np.uint32), ("c_val", np.float64), ("dense_off", np.uint64), ("dense", np.float64),
           ("l_row_off", np.uint64), ("l_ptr", np.uint64), ("l_idx", np.uint32), ("l_val", np.float64),
           ("piv_off", np.uint64), ("piv", np.float64)]


def parse_precond(b: bytes) -> dict:
    """Decode the library's preconditioner wire format (magic WXPC, version 1)."""
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
    """The exported Schwarz data in the oracle's layout: dense kept blocks (the pseudo-inverse
    S^+) become oracle kind 2, sparse LDL' factors kind 1."""
    p = parse_precond(b)
    assert p["kind"] == 1
    kinds = np.where(p["fac_kind"] == 0, 2, 1).astype(np.int32)
    nd = len(p["dense"])
    fac_val = np.concatenate([p["dense"], p["l_val"]])
    fac_ptr = p["l_ptr"].astype(np.uint64) + np.uint64(nd)
    fac_off = np.where(p["fac_kind"] == 0, p["dense_off"], p["l_row_off"]).astype(np.uint64)
    arrays = dict(slot_start=p["slot_start"], ne=p["ne"], nk=p["nk"], glob=p["glob"], pw=p["pw"],
                  inv_de=p["inv_de"], c_ptr=p["c_ptr"], c_idx=p["c_idx"], c_val=p["c_val"], fac_kind=kinds,
                  fac_off=fac_off, fac_val=fac_val, fac_ptr=fac_ptr, fac_idx=np.concatenate([np.zeros(nd, np.uint32), p["l_idx"]]), piv_off=p["piv_off"],
                  piv=p["piv"])
    return OracleSchwarz(lib, p["n"], arrays)
