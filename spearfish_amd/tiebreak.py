"""The tie-break's segment alignments (crates/map/src/lib.rs:1307-1492, the `seg` closure of
`Mapper::segmented`) for many reads at once, over a device-resident copy of the index's packed
transcript sequence (crates/index/src/lib.rs:45-151).

`PackedStore` holds the packed words on the device; `PackedStore.sequence_eq` is
`PackedRange::sequence_eq` (index lib.rs:135-137) for many range pairs.  `segments` scores every
candidate of every segment the way `seg` does with `tb.affine` set (and `affine_within` unset):
equal texts of a segment are found on the device and aligned once, left read ends reversed, Prefix
ends with `clip_ends` through align.rs:903-949 and the rest through align.rs:854-861.  All of it
goes through the C ABI (include/sf_align.h: sf_store_*, sf_tiebreak_segments).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import NamedTuple, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .align import Affine, Mode, _codes

USIZE_MAX = 2**64 - 1


def pack(codes) -> np.ndarray:
    """The index's packing: base i at bits 2 (i % 32) of u64 word i / 32 (codes masked with 3)."""
    c = _codes(codes).astype(np.uint64) & np.uint64(3)
    w = np.zeros((len(c) + 31) // 32, np.uint64)
    for k in range(32):
        part = c[k::32]
        w[: len(part)] |= part << np.uint64(2 * k)
    return w


class PackedStore:
    """Packed transcript bases on the device; candidate texts are (offset, length) ranges of it."""

    def __init__(self, words, bases: int):
        self._lib = N.load()
        w = np.ascontiguousarray(words, dtype=np.uint64)
        if len(w) < (bases + 31) // 32:
            raise ValueError("fewer words than the bases need")
        h = C.c_void_p()
        N.check(self._lib.sf_store_create(w.ctypes.data, bases, C.byref(h)))
        self._h = h
        self.bases = bases

    @classmethod
    def from_codes(cls, codes) -> "PackedStore":
        c = _codes(codes)
        return cls(pack(c), len(c))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sf_store_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def sequence_eq(self, a_off, b_off, length) -> np.ndarray:
        """index lib.rs:135-137 for ranges of equal length: bool per (a_off[i], b_off[i], length[i])."""
        a = np.ascontiguousarray(a_off, dtype=np.uint64)
        b = np.ascontiguousarray(b_off, dtype=np.uint64)
        n_ = np.ascontiguousarray(np.broadcast_to(np.asarray(length, np.uint32), a.shape))
        eq = np.empty(max(len(a), 1), np.uint8)
        N.check(self._lib.sf_store_codes_eq(self._h, a.ctypes.data, b.ctypes.data, n_.ctypes.data, len(a),
                                            eq.ctypes.data))
        return eq[: len(a)].astype(bool)


@dataclass(frozen=True)
class TieBreak:
    """The crates/map/src/lib.rs:27-214 `TieBreak` fields the segment alignments read (defaults as
    lib.rs:222-266)."""

    affine: Affine = Affine.ONT
    band: int = 16
    end_bonus: int = 5
    xdrop: int = 50
    clip_ends: bool = True
    end_diagonal: bool = False
    flank_min: int = 0
    end_bonus_cut: bool = True
    pad: int = 20
    pad_div: int = 10

    def window(self, x: int) -> int:
        """lib.rs:216-219: transcript bases searched for an unanchored read end of x bases."""
        return x + self.pad + x // self.pad_div


class Segment(NamedTuple):
    """One `seg` call: read bases [read_off, read_off + read_len) against candidate texts (store
    (offset, length) ranges); `rev` for a left read end; `cut` None or text bases."""

    read_off: int
    read_len: int
    cands: Sequence[Tuple[int, int]]
    mode: Mode
    rev: bool = False
    cut: Optional[int] = None


def segments(store: PackedStore, reads, segs: Sequence[Segment], tb: TieBreak = TieBreak()):
    """Every candidate's (whole, cut, text bases) as `seg` computes them before negating into
    deltas (lib.rs:1470-1487), and `rep`: each candidate's first equal candidate in its segment.
    Returns four numpy arrays in candidate order (segments in order, their candidates in order)."""
    lib = N.load()
    r = _codes(reads)
    arr = (N.TbSegment * max(len(segs), 1))()
    off, ln = [], []
    for i, s in enumerate(segs):
        ext = s.mode == Mode.Prefix and tb.clip_ends
        flags = (N.SEG_REVERSE if s.rev else 0) | (N.SEG_EXTEND if ext else 0)
        flags |= (N.SEG_DIAGONAL if tb.end_diagonal else 0) | (N.SEG_TRACK_CUT if tb.flank_min > 0 else 0)
        flags |= N.SEG_BONUS_WITHIN_CUT if tb.end_bonus_cut else 0
        arr[i] = N.TbSegment(s.read_off, s.read_len, len(off), len(s.cands), int(s.mode), flags,
                             tb.window(s.read_len), USIZE_MAX if s.cut is None else s.cut)
        for o, n in s.cands:
            off.append(o)
            ln.append(n)
    n = len(off)
    co = np.asarray(off + [0], np.uint64)
    cl = np.asarray(ln + [0], np.uint32)
    out = np.empty(max(n, 1), np.dtype([("whole", np.int32), ("cut", np.int32), ("cols", np.uint32)]))
    rep = np.empty(max(n, 1), np.uint32)
    p = N.TbParams(tb.affine.c(), tb.band, tb.end_bonus, tb.xdrop)
    N.check(lib.sf_tiebreak_segments(store._h, r.ctypes.data if len(r) else None, len(r), arr, len(segs),
                                     co.ctypes.data, cl.ctypes.data, n, C.byref(p), out.ctypes.data, rep.ctypes.data))
    return out["whole"][:n], out["cut"][:n], out["cols"][:n], rep[:n]
