"""Host-side mirror of spearfish's crates/map/src/align.rs interface, computed on the MI355X.

Names, argument meaning and results follow the reference (line citations per function); each
call goes through the C ABI of libsf_align.so (include/sf_align.h) to the HIP kernels.  The
reference-shaped single-pair calls are correct but launch-bound; the batch forms (PairSet,
myers_batch, affine_batch, Batch) are the ones that pay off on the device.

Sequences are base codes 0..3 (A, C, G, T; crates/lcp-core/src/encode.rs), given as bytes,
lists or numpy uint8 arrays; codes are masked with `& 3` as the reference does.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass
from typing import NamedTuple, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

NEG = -(2**31) // 4  # align.rs:693
USIZE_MAX = 2**64 - 1  # `band = usize::MAX`: exact
U32_MAX = 2**32 - 1


class Mode(enum.IntEnum):
    """align.rs:40-45."""

    Infix = 0
    Prefix = 1
    Global = 2


@dataclass(frozen=True)
class Affine:
    """align.rs:660-691: +mat per match, -mis per mismatch, -(open + ext*l) per gap of l."""

    mat: int = 2
    mis: int = 4
    open: int = 4
    ext: int = 2

    def c(self) -> N.Scoring:
        return N.Scoring(self.mat, self.mis, self.open, self.ext)


Affine.ONT = Affine(2, 4, 4, 2)  # align.rs:677-683
Affine.HIFI = Affine(1, 4, 6, 2)  # align.rs:684-690


def _codes(s) -> np.ndarray:
    if isinstance(s, np.ndarray):
        return np.ascontiguousarray(s, dtype=np.uint8)
    if isinstance(s, (bytes, bytearray)):
        return np.frombuffer(bytes(s), dtype=np.uint8)
    return np.asarray(list(s), dtype=np.uint8)


class PairSet:
    """A batch of (pattern, text) pairs as concatenated code arrays with offsets (sf_pairs)."""

    def __init__(self, pat: np.ndarray, pat_off: np.ndarray, pat_len: np.ndarray, txt: np.ndarray,
                 txt_off: np.ndarray, txt_len: np.ndarray, packed: bool = False):
        self.pat = np.ascontiguousarray(pat)
        self.pat_off = np.ascontiguousarray(pat_off, dtype=np.uint64)
        self.pat_len = np.ascontiguousarray(pat_len, dtype=np.uint32)
        self.txt = np.ascontiguousarray(txt)
        self.txt_off = np.ascontiguousarray(txt_off, dtype=np.uint64)
        self.txt_len = np.ascontiguousarray(txt_len, dtype=np.uint32)
        self.packed = packed
        assert len(self.pat_off) == len(self.pat_len) == len(self.txt_off) == len(self.txt_len)

    @classmethod
    def from_seqs(cls, patterns: Sequence, texts: Sequence) -> "PairSet":
        ps = [_codes(p) for p in patterns]
        ts = [_codes(t) for t in texts]
        assert len(ps) == len(ts)
        plen = np.array([len(p) for p in ps], dtype=np.uint32)
        tlen = np.array([len(t) for t in ts], dtype=np.uint32)
        poff = np.zeros(len(ps), np.uint64)
        toff = np.zeros(len(ts), np.uint64)
        if len(ps):
            poff[1:] = np.cumsum(plen[:-1], dtype=np.uint64)
            toff[1:] = np.cumsum(tlen[:-1], dtype=np.uint64)
        pat = np.concatenate(ps) if ps and plen.sum() else np.zeros(1, np.uint8)
        txt = np.concatenate(ts) if ts and tlen.sum() else np.zeros(1, np.uint8)
        return cls(pat, poff, plen, txt, toff, tlen)

    def __len__(self) -> int:
        return len(self.pat_len)

    def pattern(self, i: int) -> np.ndarray:
        o, l = int(self.pat_off[i]), int(self.pat_len[i])
        return self.pat[o:o + l]

    def text(self, i: int) -> np.ndarray:
        o, l = int(self.txt_off[i]), int(self.txt_len[i])
        return self.txt[o:o + l]

    def c(self) -> N.Pairs:
        bases = (lambda a: a.size * 16) if self.packed else (lambda a: a.size)
        return N.Pairs(self.pat.ctypes.data, self.pat_off.ctypes.data, self.pat_len.ctypes.data,
                       self.txt.ctypes.data, self.txt_off.ctypes.data, self.txt_len.ctypes.data,
                       len(self), int(self.packed), bases(self.pat), bases(self.txt))


def _bands(band, n: int) -> np.ndarray:
    if band is None:
        band = USIZE_MAX
    return np.ascontiguousarray(np.broadcast_to(np.asarray(band, dtype=np.uint64), (n,)))


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------------------------
# batch API
# ---------------------------------------------------------------------------------------------

def myers_batch(pairs: PairSet, mode: Mode, k: int = U32_MAX, locate: int = 0):
    """dist (U32_MAX = None), and with locate >= 1 end, >= 2 start, as numpy uint32 arrays.
    Batch form of align.rs:87-200 (`distance_with` per pair); end/start are this project's."""
    lib = N.load()
    n = len(pairs)
    dist = np.empty(n, np.uint32)
    end = np.empty(n, np.uint32) if locate >= 1 else None
    start = np.empty(n, np.uint32) if locate >= 2 else None
    cp = pairs.c()
    N.check(lib.sf_myers(C.byref(cp), int(mode), k, _p(dist), _p(end), _p(start)))
    return dist, end, start


def affine_batch(pairs: PairSet, mode: Mode, sc: Affine = Affine.ONT, band=None, cigar: bool = False):
    """Batch form of align.rs:699-701 `affine`: dict with score (int32), score_col (uint32) and,
    with cigar, tstart (uint32), cigar (uint32 BAM ops) and cigar_off (uint64, n + 1)."""
    lib = N.load()
    n = len(pairs)
    bands = _bands(band, n)
    score = np.empty(n, np.int32)
    col = np.empty(n, np.uint32)
    cp, scc = pairs.c(), sc.c()
    if not cigar:
        N.check(lib.sf_affine(C.byref(cp), int(mode), C.byref(scc), bands.ctypes.data, score.ctypes.data,
                              col.ctypes.data))
        return {"score": score, "score_col": col}
    tstart = np.empty(n, np.uint32)
    cap = int(pairs.pat_len.sum(dtype=np.uint64) + pairs.txt_len.sum(dtype=np.uint64)) + n
    cig = np.empty(max(cap, 1), np.uint32)
    off = np.empty(n + 1, np.uint64)
    N.check(lib.sf_affine_cigar(C.byref(cp), int(mode), C.byref(scc), bands.ctypes.data, score.ctypes.data,
                                col.ctypes.data, tstart.ctypes.data, cig.ctypes.data, cap, off.ctypes.data))
    return {"score": score, "score_col": col, "tstart": tstart, "cigar": cig[:int(off[-1]) if n else 0],
            "cigar_off": off}


class Batch:
    """A device-resident batch (sf_batch): packed into HBM once, run many times."""

    def __init__(self, pairs: PairSet):
        self.lib = N.load()
        self.n = len(pairs)
        self._pairs = pairs
        h = C.c_void_p()
        cp = pairs.c()
        N.check(self.lib.sf_batch_create(C.byref(cp), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.sf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def myers(self, mode: Mode, k: int = U32_MAX, locate: int = 0):
        N.check(self.lib.sf_batch_myers(self.h, int(mode), k, locate))

    def affine(self, mode: Mode, sc: Affine = Affine.ONT, band=None, cigar: bool = False, bands=None):
        if bands is None:
            bands = _bands(band, self.n)
        self._bands = bands
        scc = sc.c()
        N.check(self.lib.sf_batch_affine(self.h, int(mode), C.byref(scc), bands.ctypes.data, int(cigar)))

    def use_located(self, on: bool = True):
        """Restrict texts to the windows [start, end) of the last myers(locate=2) run (or restore)."""
        N.check(self.lib.sf_batch_use_located(self.h, int(on)))

    def sync(self):
        N.check(self.lib.sf_batch_sync(self.h))

    def myers_results(self, locate: int = 0):
        dist = np.empty(self.n, np.uint32)
        end = np.empty(self.n, np.uint32) if locate >= 1 else None
        start = np.empty(self.n, np.uint32) if locate >= 2 else None
        N.check(self.lib.sf_batch_myers_results(self.h, _p(dist), _p(end), _p(start)))
        return dist, end, start

    def affine_results(self, cigar: bool = False):
        score = np.empty(self.n, np.int32)
        col = np.empty(self.n, np.uint32)
        if not cigar:
            N.check(self.lib.sf_batch_affine_results(self.h, _p(score), _p(col), None, None, 0, None))
            return {"score": score, "score_col": col}
        tstart = np.empty(self.n, np.uint32)
        off = np.empty(self.n + 1, np.uint64)
        N.check(self.lib.sf_batch_affine_results(self.h, _p(score), _p(col), _p(tstart), None, 0, _p(off)))
        cig = np.empty(max(int(off[-1]), 1), np.uint32)
        N.check(self.lib.sf_batch_affine_results(self.h, None, None, None, _p(cig), len(cig), _p(off)))
        return {"score": score, "score_col": col, "tstart": tstart, "cigar": cig[:int(off[-1])], "cigar_off": off}

    def autotune(self, mode: Mode, sc: Affine = Affine.ONT, band=None, bands=None):
        """Measure on the device the score kernels' shape classes this batch's pairs could take
        (sf_batch_autotune); later score runs plan with the measurements."""
        if bands is None:
            bands = _bands(band, self.n)
        scc = sc.c()
        N.check(self.lib.sf_batch_autotune(self.h, int(mode), C.byref(scc), bands.ctypes.data))

    def capture_begin(self):
        """Record the following myers() and score-only affine() calls as one HIP graph."""
        N.check(self.lib.sf_batch_capture_begin(self.h))

    def capture_end(self):
        N.check(self.lib.sf_batch_capture_end(self.h))

    def replay(self, times: int = 1):
        """Launch the captured step `times` times (asynchronous; results as after those calls)."""
        N.check(self.lib.sf_batch_replay(self.h, times))

    def timing(self) -> N.Timing:
        t = N.Timing()
        N.check(self.lib.sf_batch_timing(self.h, C.byref(t)))
        return t

    def reset_timing(self):
        self.lib.sf_batch_reset_timing(self.h)


# ---------------------------------------------------------------------------------------------
# passes with per-row bookkeeping: X-drop extension, cut, band layout, local bands
# ---------------------------------------------------------------------------------------------

class Dp(NamedTuple):
    """align.rs:2111-2129 `Dp` (clip_cell as (clip_i, clip_j))."""

    score: int
    score_col: int
    clip_best: int
    clip_i: int
    clip_j: int
    text_read: int
    last_row: int
    cut_best: int
    cut_score: int


def _u64s(v, n):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(USIZE_MAX if v is None else v, dtype=np.uint64), (n,)))


def affine_dp_batch(pairs: PairSet, mode: Mode, sc: Affine = Affine.ONT, band=None, geom=None,
                    clip: Optional[Tuple[int, bool]] = None, cut=None) -> list:
    """align.rs:2183-2208 `affine_dp` for every pair: band, geom and cut per pair or shared
    (geom None: each text's length; cut None: usize::MAX); clip = (xdrop, clip_cell) or None."""
    lib = N.load()
    n = len(pairs)
    out = (N.Dp * max(n, 1))()
    bands = _bands(band, n)
    g = None if geom is None else _u64s(geom, n)
    c = None if cut is None else _u64s(cut, n)
    has, xd, cell = (0, 0, 0) if clip is None else (1, int(clip[0]), int(bool(clip[1])))
    cp, scc = pairs.c(), sc.c()
    N.check(lib.sf_affine_dp(C.byref(cp), int(mode), C.byref(scc), bands.ctypes.data, _p(g), has, xd, cell, _p(c),
                             C.cast(out, C.c_void_p)))
    return [Dp(o.score, o.score_col, o.clip_best, o.clip_i, o.clip_j, o.text_read, o.last_row, o.cut_best,
               o.cut_score) for o in out[:n]]


def affine_dp(pattern, text, mode: Mode, sc: Affine = Affine.ONT, band: int = USIZE_MAX, clip=None,
              geom: Optional[int] = None, cut: int = USIZE_MAX) -> Dp:
    """align.rs:2183-2208 `affine_dp` (clip = (xdrop, clip_cell) or None; geom None: len(text))."""
    return affine_dp_batch(PairSet.from_seqs([pattern], [text]), mode, sc, band, geom, clip, cut)[0]


def affine_extend(pattern, text, sc: Affine, band: int, end_bonus: int, xdrop: int) -> int:
    """align.rs:710-729."""
    lib = N.load()
    ps = PairSet.from_seqs([pattern], [text])
    out = np.empty(1, np.int32)
    cp, scc = ps.c(), sc.c()
    N.check(lib.sf_affine_extend(C.byref(cp), C.byref(scc), _bands(band, 1).ctypes.data, end_bonus, xdrop,
                                 out.ctypes.data))
    return int(out[0])


def affine_extend_window(pattern, text, window: int, sc: Affine, band: int, end_bonus: int, xdrop: int):
    """align.rs:957-982: (score, pattern bases, text bases, text bases read)."""
    lib = N.load()
    ps = PairSet.from_seqs([pattern], [text])
    out = N.ExtendWindow()
    cp, scc = ps.c(), sc.c()
    N.check(lib.sf_affine_extend_window(C.byref(cp), _u64s(window, 1).ctypes.data, C.byref(scc),
                                        _bands(band, 1).ctypes.data, end_bonus, xdrop, C.byref(out)))
    return (out.score, out.pat_bases, out.txt_bases, out.txt_read)


def affine_extend_sorted_cut_bonus(pattern, texts: Sequence, sc: Affine, band: int, end_bonus: int, xdrop: int,
                                   diagonal: bool, window: int, cut: int, track_cut: bool,
                                   bonus_within_cut: bool) -> list:
    """align.rs:903-949: (whole, cut, text bases covered) per text; equal to the sorted call."""
    if not len(texts):
        return []
    lib = N.load()
    n = len(texts)
    ps = PairSet.from_seqs([pattern] * n, texts)
    out = (N.ExtendCut * n)()
    cp, scc = ps.c(), sc.c()
    N.check(lib.sf_affine_extend_cut_bonus(C.byref(cp), C.byref(scc), _bands(band, n).ctypes.data, end_bonus, xdrop,
                                           int(diagonal), _u64s(window, n).ctypes.data, _u64s(cut, n).ctypes.data,
                                           int(track_cut), int(bonus_within_cut), C.cast(out, C.c_void_p)))
    return [(o.whole, o.cut, o.cols) for o in out]


def affine_extend_sorted_cut(pattern, texts: Sequence, sc: Affine, band: int, end_bonus: int, xdrop: int,
                             diagonal: bool, window: int, cut: int) -> list:
    """align.rs:880-901."""
    return affine_extend_sorted_cut_bonus(pattern, texts, sc, band, end_bonus, xdrop, diagonal, window, cut, True,
                                          False)


def local_band_ends_batch(pairs: PairSet, sc: Affine, dlo, dhi, end_bonus: Tuple[int, int]) -> np.ndarray:
    """align.rs:741-760 for every pair (dlo, dhi per pair or shared)."""
    lib = N.load()
    n = len(pairs)
    lo = np.ascontiguousarray(np.broadcast_to(np.asarray(dlo, np.int64), (n,)))
    hi = np.ascontiguousarray(np.broadcast_to(np.asarray(dhi, np.int64), (n,)))
    out = np.empty(max(n, 1), np.int32)
    cp, scc = pairs.c(), sc.c()
    N.check(lib.sf_local_band_ends(C.byref(cp), C.byref(scc), lo.ctypes.data, hi.ctypes.data, int(end_bonus[0]),
                                   int(end_bonus[1]), out.ctypes.data))
    return out[:n]


def local_band_ends(read, text, sc: Affine, dlo: int, dhi: int, end_bonus: Tuple[int, int]) -> int:
    """align.rs:741-760: best local score over diagonals dlo..=dhi; read codes >= 4 match nothing."""
    return int(local_band_ends_batch(PairSet.from_seqs([read], [text]), sc, dlo, dhi, end_bonus)[0])


def local_band(read, text, sc: Affine, dlo: int, dhi: int, end_bonus: int) -> int:
    """align.rs:735-737."""
    return local_band_ends(read, text, sc, dlo, dhi, (end_bonus, end_bonus))


# ---------------------------------------------------------------------------------------------
# the reference's single-pair interface (align.rs), one pair per device call
# ---------------------------------------------------------------------------------------------

class Pattern:
    """align.rs:68-85 `Pattern`: a pattern reused across texts (its match masks are built on the
    device, per launch)."""

    def __init__(self, pattern):
        self.codes = _codes(pattern)
        self.m = len(self.codes)


def distance_with(pat: Pattern, text, mode: Mode, k: int) -> Optional[int]:
    """align.rs:92-168: the edit distance in `mode`, or None if it exceeds k."""
    d, _, _ = myers_batch(PairSet.from_seqs([pat.codes], [text]), mode, k)
    return None if d[0] == U32_MAX else int(d[0])


def distance(pattern, text, mode: Mode, k: int) -> Optional[int]:
    """align.rs:87-90."""
    return distance_with(Pattern(pattern), text, mode, k)


def batch_distances(pat: Pattern, texts: Sequence, mode: Mode, k: int) -> list:
    """align.rs:170-200: the same results as one distance_with per text."""
    if not len(texts):
        return []
    d, _, _ = myers_batch(PairSet.from_seqs([pat.codes] * len(texts), texts), mode, k)
    return [None if x == U32_MAX else int(x) for x in d]


def banded_batch(pairs: PairSet, group_off, mode: Mode, cap: int, distinct: bool = False) -> np.ndarray:
    """crates/map/src/lib.rs:851-911 `Mapper::banded` for many reads: group g is pairs
    [group_off[g], group_off[g + 1]), one read against its texts in order; one u32 per pair."""
    lib = N.load()
    n = len(pairs)
    go = np.ascontiguousarray(group_off, dtype=np.uint64)
    out = np.empty(max(n, 1), np.uint32)
    cp = pairs.c()
    N.check(lib.sf_myers_banded(C.byref(cp), go.ctypes.data, len(go) - 1, int(mode), cap, int(distinct),
                                out.ctypes.data))
    return out[:n]


def banded(read, texts: Sequence, mode: Mode, cap: int, distinct: bool = False) -> list:
    """crates/map/src/lib.rs:851-911 `Mapper::banded`: one read against its texts."""
    if not len(texts):
        return []
    ps = PairSet.from_seqs([read] * len(texts), texts)
    return [int(x) for x in banded_batch(ps, [0, len(texts)], mode, cap, distinct)]


def infix_distance(pattern, text) -> int:
    """align.rs:655-658."""
    return distance(pattern, text, Mode.Infix, U32_MAX)


def affine(pattern, text, mode: Mode, sc: Affine = Affine.ONT, band: int = USIZE_MAX) -> int:
    """align.rs:699-701: the banded affine score (a lower bound; exact with usize::MAX)."""
    r = affine_batch(PairSet.from_seqs([pattern], [text]), mode, sc, band)
    return int(r["score"][0])


def affine_sorted(pattern, texts: Sequence, mode: Mode, sc: Affine = Affine.ONT, band: int = USIZE_MAX) -> list:
    """align.rs:854-861: equal to separate `affine` calls (align.rs:2210-2213)."""
    if not len(texts):
        return []
    r = affine_batch(PairSet.from_seqs([pattern] * len(texts), texts), mode, sc, band)
    return [int(x) for x in r["score"]]
