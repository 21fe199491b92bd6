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


# sf_tb_segment as a numpy record (the same layout as the C struct).
SEGMENT = np.dtype([("read_off", np.uint64), ("read_len", np.uint32), ("cand0", np.uint32), ("ncand", np.uint32),
                    ("mode", np.int32), ("flags", np.uint32), ("window", np.uint64), ("cut", np.uint64)], align=True)
RESULT = np.dtype([("whole", np.int32), ("cut", np.int32), ("cols", np.uint32)])


def flags(mode: Mode, rev: bool, tb: TieBreak) -> int:
    """The SF_SEG_* flags of a `seg` call on a `mode` segment under `tb` (lib.rs:1446-1463)."""
    f = N.SEG_REVERSE if rev else 0
    if mode == Mode.Prefix and tb.clip_ends:
        f |= N.SEG_EXTEND
    f |= (N.SEG_DIAGONAL if tb.end_diagonal else 0) | (N.SEG_TRACK_CUT if tb.flank_min > 0 else 0)
    return f | (N.SEG_BONUS_WITHIN_CUT if tb.end_bonus_cut else 0)


def segments_arrays(store: PackedStore, reads, segs: np.ndarray, cand_off, cand_len, tb: TieBreak = TieBreak()):
    """sf_tiebreak_segments on arrays: segs a SEGMENT array (its `flags` and `window` as set by the
    caller), candidates (store offset, length); returns (RESULT array, rep)."""
    lib = N.load()
    r = _codes(reads)
    sg = np.ascontiguousarray(segs, SEGMENT)
    co = np.ascontiguousarray(cand_off, np.uint64)
    cl = np.ascontiguousarray(cand_len, np.uint32)
    n = len(co)
    out = np.empty(max(n, 1), RESULT)
    rep = np.empty(max(n, 1), np.uint32)
    p = N.TbParams(tb.affine.c(), tb.band, tb.end_bonus, tb.xdrop)
    N.check(lib.sf_tiebreak_segments(store._h, r.ctypes.data if len(r) else None, len(r), sg.ctypes.data, len(sg),
                                     co.ctypes.data, cl.ctypes.data, n, C.byref(p), out.ctypes.data, rep.ctypes.data))
    return out[:n], rep[:n]


class TbBatch:
    """A tie-break call kept on the device (sf_tb_batch_*): made once from its inputs, then run and
    read as often as wanted, with any TieBreak scoring parameters."""

    def __init__(self, store: PackedStore, reads, segs: np.ndarray, cand_off, cand_len):
        self._lib = N.load()
        r = _codes(reads)
        sg = np.ascontiguousarray(segs, SEGMENT)
        co = np.ascontiguousarray(cand_off, np.uint64)
        cl = np.ascontiguousarray(cand_len, np.uint32)
        self.n = len(co)
        h = C.c_void_p()
        N.check(self._lib.sf_tb_batch_create(store._h, r.ctypes.data if len(r) else None, len(r), sg.ctypes.data,
                                             len(sg), co.ctypes.data, cl.ctypes.data, self.n, C.byref(h)))
        self._h = h

    def run(self, tb: TieBreak = TieBreak()):
        p = N.TbParams(tb.affine.c(), tb.band, tb.end_bonus, tb.xdrop)
        N.check(self._lib.sf_tb_batch_run(self._h, C.byref(p)))

    def results(self):
        out = np.empty(max(self.n, 1), RESULT)
        rep = np.empty(max(self.n, 1), np.uint32)
        N.check(self._lib.sf_tb_batch_results(self._h, out.ctypes.data, rep.ctypes.data))
        return out[: self.n], rep[: self.n]

    def timing(self, reset: bool = False) -> N.TbTiming:
        t = N.TbTiming()
        N.check(self._lib.sf_tb_batch_timing(self._h, C.byref(t), int(reset)))
        return t

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sf_tb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def segments(store: PackedStore, reads, segs: Sequence[Segment], tb: TieBreak = TieBreak()):
    """Every candidate's (whole, cut, text bases) as `seg` computes them before negating into
    deltas (lib.rs:1470-1487), and `rep`: each candidate's first equal candidate in its segment.
    Returns four numpy arrays in candidate order (segments in order, their candidates in order)."""
    arr = np.zeros(len(segs), SEGMENT)
    off, ln = [], []
    for i, s in enumerate(segs):
        arr[i] = (s.read_off, s.read_len, len(off), len(s.cands), int(s.mode), flags(s.mode, s.rev, tb),
                  tb.window(s.read_len), USIZE_MAX if s.cut is None else s.cut)
        off += [o for o, _ in s.cands]
        ln += [n for _, n in s.cands]
    out, rep = segments_arrays(store, reads, arr, off, ln, tb)
    return out["whole"], out["cut"], out["cols"], rep


def calls_from_isoform_segments(tx_off, tx_len, reads, tb: TieBreak = TieBreak()):
    """The tie-break's calls for reads whose isoform segments are known (the serving workload's
    `Reads`: per read, consecutive segments with the same read slice are that slice against each
    isoform), shaped as `segmented` makes them (lib.rs:1496-1672): a read's first slice is its left
    end (reversed Prefix against windows of tb.window(slice length) bases ending where the slice's
    texts end), its last slice its right end (Prefix against windows starting where the texts
    start), the ones between interior stretches (Global against the texts).  An end's cut is the
    shortest window that stops at its transcript's end, where longer ones exist (lib.rs:1502-1510).
    Returns (SEGMENT array, candidate offsets, candidate lengths); reads one-slice long give a right
    end."""
    tx_off = np.asarray(tx_off, np.uint64)
    tx_end = tx_off + np.asarray(tx_len, np.uint64)
    so = reads.seg_read_off.astype(np.uint64)
    sl = reads.seg_read_len.astype(np.uint64)
    to = reads.seg_tx_off.astype(np.uint64)
    tl = reads.seg_tx_len.astype(np.uint64)
    tx = np.searchsorted(tx_off, to, side="right") - 1  # each segment's transcript
    segs, coff, clen = [], [], []
    for i in range(len(reads)):
        s0, ns = int(reads.seg0[i]), int(reads.nseg[i])
        groups, k = [], s0
        while k < s0 + ns:  # consecutive segments of one read slice
            e = k + 1
            while e < s0 + ns and so[e] == so[k] and sl[e] == sl[k]:
                e += 1
            groups.append((k, e))
            k = e
        for gi, (a, e) in enumerate(groups):
            left, right = gi == 0 and len(groups) > 1, gi == len(groups) - 1
            m = int(sl[a])
            rev = left
            mode = Mode.Prefix if (left or right) else Mode.Global
            w = tb.window(m)
            cands, short = [], []
            for q in range(a, e):
                t = int(tx[q])
                if left:  # the window ends where the text ends
                    hi = int(to[q] + tl[q])
                    lo = max(int(tx_off[t]), hi - w)
                    short.append(hi - lo < w)
                elif right:  # the window starts where the text starts
                    lo = int(to[q])
                    hi = min(int(tx_end[t]), lo + w)
                    short.append(hi - lo < w)
                else:
                    lo, hi = int(to[q]), int(to[q] + tl[q])
                    short.append(False)
                cands.append((lo, hi - lo))
            cut = USIZE_MAX
            lens = [n for _, n in cands]
            sh = [n for n, s in zip(lens, short) if s]
            if sh and (tb.flank_min > 0 or tb.end_bonus_cut) and any(n > min(sh) for n in lens):
                cut = min(sh)
            segs.append((int(reads.read_off[i] + so[a]), m, len(coff), len(cands), int(mode), flags(mode, rev, tb),
                         w, cut))
            coff += [c[0] for c in cands]
            clen += [c[1] for c in cands]
    return np.array(segs, SEGMENT), np.array(coff, np.uint64), np.array(clen, np.uint32)
