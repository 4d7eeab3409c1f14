"""Reader for the reference's serialized preconditioner, postcard wire format v13.

The reference pickles a ``Preconditioner`` as postcard bytes of its serde layout
(crates/within-py/src/config.rs:367-388, crates/within/src/operator/schwarz.rs:116-128):

    Preconditioner = Variant                        (serde transparent)
    Variant        = 0 Additive(SchwarzPreconditioner) | 1 Diagonal{inv_diag: [f64]}
    SchwarzPreconditioner = {subdomains: [SubdomainEntry], n_dofs: usize}
                                                    (schwarz-precond/src/schwarz/preconditioner.rs:20-32)
    SubdomainEntry = {core: SubdomainCore, solver: BlockElimSolver}     (local_solve.rs:150-166)
    SubdomainCore  = {global_indices: [u32], partition_weights: PartitionWeights}  (domain.rs:191-204)
    PartitionWeights = 0 Uniform(usize) | 1 NonUniform([f64])           (domain.rs:12-20)
    BlockElimSolver = {cross_tab: {c: CsrBlock, ct: CsrBlock}, inv_diag_elim: [f64],
                       reduced_factor: ReducedFactor, coordinates: CoordinateMap}
                                                    (block_elim/solver.rs:80-97, domain/cross_tab.rs:95-101)
    CsrBlock       = {indptr: [u32], indices: [u32], data: [f64], nrows: usize, ncols: usize}
                                                    (csr_block.rs:15-22)
    ReducedFactor  = 0 Direct{factor: Factor, grounding: 0 Floating | 1 Grounded}
                   | 1 Cover{inner: Factor, m: usize}                    (block_elim/factor.rs:15-31)
    CoordinateMap  = 0 Canonical | 1 Scaled([f64])  (domain/factor_pairs/sddm.rs:50-58)

Postcard writes unsigned integers as LEB128 varints, enum variants as varint indices in
declaration order, sequences as a varint length followed by the elements, and f64 as eight
little-endian bytes.

``Factor`` belongs to the third-party approx-chol crate, whose source is not available here; its
bytes are kept opaque.  Every such factor begins with approx-chol's format tag, the varint
0x41430003 (bytes 83 80 8c 8a 04; a v12 payload carries 0x41430002), and its end is located as
the position after which the rest of the payload decodes as the remainder of a valid
preconditioner: the factor's trailing grounding or cover size, the coordinate map, then either the
next subdomain or the final n_dofs.  The reference's own deserializer checks (CSR structure, the
transpose's shape, the eliminated diagonal's length, the scaled map's length, n_dofs covering
every subdomain index; block_elim/solver.rs:98-152, preconditioner.rs:34-66) are applied on the
way, so the located end is the one that yields a preconditioner the reference would accept.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

FACTOR_TAG = 0x41430003          # approx-chol Factor format tag of wire v13
FACTOR_TAG_V12 = 0x41430002


class WireError(ValueError):
    """The bytes are not a valid v13 preconditioner payload."""


@dataclass
class Csr:
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    nrows: int
    ncols: int

    def dense(self) -> np.ndarray:
        out = np.zeros((self.nrows, self.ncols))
        for i in range(self.nrows):
            s, e = int(self.indptr[i]), int(self.indptr[i + 1])
            out[i, self.indices[s:e]] = self.data[s:e]
        return out


@dataclass
class Subdomain:
    global_indices: np.ndarray          # eliminated block first, then the kept block
    partition_weights: np.ndarray       # one per global index (1.0 when Uniform)
    uniform: bool
    c: Csr                              # eliminated x kept cross block
    ct: Csr
    inv_diag_elim: np.ndarray
    reduced_factor: str                 # "Direct" or "Cover"
    grounding: Optional[str]            # Direct: "Floating" / "Grounded"
    cover_m: Optional[int]              # Cover: the signed reduced dimension
    coordinates: Optional[np.ndarray]   # Scaled factors; None for Canonical
    factor_span: tuple                  # (start, end) byte offsets of the opaque Factor

    @property
    def n_eliminated(self) -> int:
        return self.c.nrows

    @property
    def n_kept(self) -> int:
        return self.c.ncols


@dataclass
class WirePreconditioner:
    variant: str                                  # "Additive" or "Diagonal"
    n_dofs: int
    subdomains: List[Subdomain] = field(default_factory=list)
    inv_diag: Optional[np.ndarray] = None         # Diagonal only


class _Reader:
    __slots__ = ("b", "pos")

    def __init__(self, b: bytes, pos: int = 0):
        self.b = b
        self.pos = pos

    def varint(self, limit_bits: int = 64) -> int:
        v, shift = 0, 0
        while True:
            if self.pos >= len(self.b):
                raise WireError("truncated varint")
            x = self.b[self.pos]
            self.pos += 1
            v |= (x & 0x7F) << shift
            if not x & 0x80:
                break
            shift += 7
            if shift >= limit_bits + 7:
                raise WireError("varint too long")
        if v >> limit_bits:
            raise WireError("varint out of range")
        return v

    def u8(self) -> int:
        if self.pos >= len(self.b):
            raise WireError("truncated")
        x = self.b[self.pos]
        self.pos += 1
        return x

    def length(self, item_bytes_min: int) -> int:
        n = self.varint()
        if n * item_bytes_min > len(self.b) - self.pos:
            raise WireError("sequence length exceeds the payload")
        return n

    def f64s(self) -> np.ndarray:
        n = self.length(8)
        a = np.frombuffer(self.b, "<f8", n, self.pos).copy()
        self.pos += 8 * n
        return a

    def u32s(self) -> np.ndarray:
        n = self.length(1)
        return np.array([self.varint(32) for _ in range(n)], dtype=np.uint32)


def _csr(r: _Reader) -> Csr:
    indptr = r.u32s()
    indices = r.u32s()
    data = r.f64s()
    nrows = r.varint()
    ncols = r.varint()
    # csr_block.rs:31-45
    if nrows + 1 != len(indptr) or len(data) != len(indices):
        raise WireError("CSR block is not structurally valid")
    if indptr[0] != 0 or int(indptr[-1]) != len(indices) or np.any(np.diff(indptr.astype(np.int64)) < 0):
        raise WireError("CSR block is not structurally valid")
    if len(indices) and int(indices.max()) >= ncols:
        raise WireError("CSR block is not structurally valid")
    return Csr(indptr, indices, data, nrows, ncols)


def _prefix(r: _Reader) -> dict:
    """A subdomain up to (and including) its factor's format tag."""
    gi = r.u32s()
    pw_kind = r.varint(32)
    if pw_kind == 0:
        count = r.varint()
        if count != len(gi):
            raise WireError("partition weight count must match index count")
        pw, uniform = np.ones(len(gi)), True
    elif pw_kind == 1:
        pw, uniform = r.f64s(), False
        if len(pw) != len(gi):
            raise WireError("partition weight count must match index count")
    else:
        raise WireError(f"unknown PartitionWeights variant {pw_kind}")
    c = _csr(r)
    ct = _csr(r)
    if ct.nrows != c.ncols or ct.ncols != c.nrows:
        raise WireError("cross_tab.ct shape disagrees with c")
    if len(gi) != c.nrows + c.ncols:  # local_solve.rs:53-61
        raise WireError("subdomain index count disagrees with its cross-tab")
    inv = r.f64s()
    if len(inv) != c.nrows:
        raise WireError("inv_diag_elim length disagrees with eliminated block")
    rf = r.varint(32)
    if rf not in (0, 1):
        raise WireError(f"unknown ReducedFactor variant {rf}")
    start = r.pos
    tag = r.varint(32)
    if tag == FACTOR_TAG_V12:
        raise WireError("approx-chol factor format 0x41430002 (wire v12); expected v13")
    if tag != FACTOR_TAG:
        raise WireError(f"not an approx-chol factor (tag {tag:#x})")
    return dict(gi=gi, pw=pw, uniform=uniform, c=c, ct=ct, inv=inv, rf=rf, start=start, after_tag=r.pos)


def _tail(r: _Reader, p: dict) -> Optional[dict]:
    """Fields after the opaque factor, or None when they do not decode here."""
    try:
        if p["rf"] == 0:
            g = r.u8()
            if g > 1:
                return None
            grounding, m = ("Floating", "Grounded")[g], None
        else:
            m = r.varint()
            if m != p["c"].ncols:  # the cover exposes the kept block (factor.rs:52-63)
                return None
            grounding = None
        kind = r.varint(32)
        if kind == 0:
            coords = None
        elif kind == 1:
            coords = r.f64s()
            if len(coords) != p["c"].nrows + p["c"].ncols:
                return None
        else:
            return None
    except WireError:
        return None
    return dict(grounding=grounding, m=m, coords=coords)


def _n_dofs_at(b: bytes, pos: int) -> Optional[int]:
    r = _Reader(b, pos)
    try:
        n = r.varint()
    except WireError:
        return None
    return n if r.pos == len(b) else None


def _subdomains(b: bytes, pos: int, count: int) -> tuple:
    """Decodes `count` subdomains from pos, then n_dofs.  A factor's end is the first position
    after its tag where its trailing fields decode and are followed by the next subdomain's
    prefix up to that subdomain's own factor tag (or, after the last one, by n_dofs ending the
    payload)."""
    subs: List[Subdomain] = []
    r = _Reader(b, pos)
    p = _prefix(r) if count else None
    for i in range(count):
        found = None
        for end in range(p["after_tag"], len(b)):
            rt = _Reader(b, end)
            t = _tail(rt, p)
            if t is None:
                continue
            if i + 1 < count:
                try:
                    nxt = _prefix(_Reader(b, rt.pos))
                except WireError:
                    continue
                found = (end, t, nxt)
            else:
                n_dofs = _n_dofs_at(b, rt.pos)
                if n_dofs is None:
                    continue
                found = (end, t, n_dofs)
            break
        if found is None:
            raise WireError(f"subdomain {i}: no end of its factor leaves a decodable remainder")
        end, t, nxt = found
        subs.append(Subdomain(p["gi"], p["pw"], p["uniform"], p["c"], p["ct"], p["inv"],
                              ("Direct", "Cover")[p["rf"]], t["grounding"], t["m"], t["coords"],
                              (p["start"], end)))
        if i + 1 < count:
            p = nxt
        else:
            return subs, nxt
    return subs, _n_dofs_at(b, r.pos)


def decode(data: bytes) -> WirePreconditioner:
    """Decodes a reference v13 preconditioner payload; raises WireError when it is not one."""
    b = bytes(data)
    r = _Reader(b)
    variant = r.varint(32)
    if variant == 1:  # Diagonal{inv_diag}
        inv = r.f64s()
        if r.pos != len(b):
            raise WireError("trailing bytes after the diagonal preconditioner")
        return WirePreconditioner("Diagonal", len(inv), inv_diag=inv)
    if variant != 0:
        raise WireError(f"unknown preconditioner variant {variant}")
    n_sub = r.length(8)
    subs, n_dofs = _subdomains(b, r.pos, n_sub)
    if n_dofs is None:
        raise WireError("trailing bytes or a missing n_dofs after the subdomains")
    span = max((int(s.global_indices.max()) + 1 for s in subs if len(s.global_indices)), default=0)
    if n_dofs < span:  # preconditioner.rs:48-60
        raise WireError(f"n_dofs ({n_dofs}) is below the covered subdomain index span ({span})")
    return WirePreconditioner("Additive", n_dofs, subs)


def is_reference_payload(data: bytes) -> bool:
    """True when the bytes decode as a reference v13 preconditioner."""
    try:
        decode(data)
        return True
    except WireError:
        return False
