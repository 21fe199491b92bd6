"""ctypes binding of the CPU oracle (libsf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, to check the product's results against.  The product package (spearfish_amd) never imports it.

The functions mirror crates/map/src/align.rs (the reference, Rust) one to one; see sf_oracle.h for
the line citations.  Sequences are bytes / bytearray / numpy uint8 arrays of base codes (0..3).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsf_oracle.so")

INFIX, PREFIX, GLOBAL = 0, 1, 2
NEG = -(2**31) // 4
NONE = 0xFFFFFFFF
UMAX = 2**64 - 1


class Scoring(C.Structure):
    _fields_ = [("mat", C.c_int32), ("mis", C.c_int32), ("open", C.c_int32), ("ext", C.c_int32)]


ONT = (2, 4, 4, 2)  # align.rs:677-683
HIFI = (1, 4, 6, 2)  # align.rs:684-690


class Dp(C.Structure):
    _fields_ = [
        ("score", C.c_int32),
        ("score_col", C.c_uint64),
        ("clip_best", C.c_int32),
        ("clip_i", C.c_uint64),
        ("clip_j", C.c_uint64),
        ("text_read", C.c_uint64),
        ("last_row", C.c_uint64),
        ("cut_best", C.c_int32),
        ("cut_score", C.c_int32),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = declare(C.CDLL(LIB_PATH))
    return _lib


def declare(L):
    """Sets the ctypes signatures of the oracle's functions on a loaded copy of the library."""
    if True:
        u8p, u64, i32, u32 = C.c_char_p, C.c_uint64, C.c_int32, C.c_uint32
        sc = C.POINTER(Scoring)
        L.sfo_distance.argtypes = [u8p, u64, u8p, u64, C.c_int, u32]
        L.sfo_distance.restype = u32
        L.sfo_batch_distances.argtypes = [u8p, u64, C.POINTER(C.c_char_p), C.POINTER(u64), u64, C.c_int, u32,
                                          C.POINTER(u32)]
        L.sfo_locate.argtypes = [u8p, u64, u8p, u64, C.c_int, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
        L.sfo_banded.argtypes = [u8p, u64, C.POINTER(C.c_char_p), C.POINTER(u64), u64, C.c_int, u32, C.c_int,
                                 C.POINTER(u32)]
        L.sfo_sequential.argtypes = [u8p, u64, C.POINTER(C.c_char_p), C.POINTER(u64), u64, C.c_int, u32,
                                     C.POINTER(u32)]
        L.sfo_affine_dp.argtypes = [u8p, u64, u8p, u64, C.c_int, sc, u64, u64, C.c_int, i32, C.c_int, u64,
                                    C.POINTER(Dp)]
        L.sfo_affine.argtypes = [u8p, u64, u8p, u64, C.c_int, sc, u64]
        L.sfo_affine.restype = i32
        L.sfo_affine_extend.argtypes = [u8p, u64, u8p, u64, sc, u64, i32, i32]
        L.sfo_affine_extend.restype = i32
        L.sfo_affine_extend_window.argtypes = [u8p, u64, u8p, u64, u64, sc, u64, i32, i32, C.POINTER(C.c_int64)]
        L.sfo_affine_extend_cut_bonus.argtypes = [u8p, u64, u8p, u64, sc, u64, i32, i32, C.c_int, u64, u64, C.c_int,
                                                  C.c_int, C.POINTER(C.c_int64)]
        L.sfo_local_band_ends.argtypes = [u8p, u64, u8p, u64, sc, C.c_int64, C.c_int64, i32, i32]
        L.sfo_local_band_ends.restype = i32
        L.sfo_align.argtypes = [u8p, u64, u8p, u64, C.c_int, sc, u64, C.POINTER(u32), u64, C.POINTER(i32),
                                C.POINTER(u64), C.POINTER(u64)]
        L.sfo_align.restype = C.c_int64
        L.sfo_naive_edit.argtypes = [u8p, u64, u8p, u64, C.c_int]
        L.sfo_naive_edit.restype = u32
        L.sfo_naive_edit_lastrow.argtypes = [u8p, u64, u8p, u64, C.c_int, C.POINTER(u32)]
        L.sfo_naive_affine.argtypes = [u8p, u64, u8p, u64, C.c_int, sc]
        L.sfo_naive_affine.restype = i32
        L.sfo_naive_extend.argtypes = [u8p, u64, u8p, u64, sc]
        L.sfo_naive_extend.restype = i32
        L.sfo_naive_local.argtypes = [u8p, u64, u8p, u64, sc, i32, i32]
        L.sfo_naive_local.restype = i32
        L.sfo_naive_local_band.argtypes = [u8p, u64, u8p, u64, sc, C.c_int64, C.c_int64, i32, i32]
        L.sfo_naive_local_band.restype = i32
        L.sfo_affine_cells.argtypes = [u64, u64, C.c_int, u64]
        L.sfo_affine_cells.restype = u64
        vp = C.c_void_p
        L.sfo_batch_myers.argtypes = [vp, vp, vp, vp, vp, vp, u64, C.c_int, u32, vp, C.c_int]
        L.sfo_batch_affine.argtypes = [vp, vp, vp, vp, vp, vp, u64, C.c_int, sc, vp, vp, vp, C.c_int]
        L.sfo_batch_affine_avx512.argtypes = [vp, vp, vp, vp, vp, vp, u64, C.c_int, sc, vp, vp, vp, C.c_int]
        L.sfo_batch_affine_avx512.restype = C.c_int
        L.sfo_avx512_available.argtypes = []
        L.sfo_avx512_available.restype = C.c_int
        L.sfo_packed_eq.argtypes = [vp, u64, u64, u64, u32]
        L.sfo_packed_eq.restype = C.c_int
        L.sfo_tiebreak_segments.argtypes = [vp, u64, vp, vp, u64, vp, vp, sc, u64, i32, i32, vp, vp, C.c_int, C.c_int]
        L.sfo_tiebreak_segments.restype = C.c_int
    return L


def _b(s) -> bytes:
    if isinstance(s, bytes):
        return s
    if isinstance(s, np.ndarray):
        return s.astype(np.uint8, copy=False).tobytes()
    return bytes(s)


def _sc(sc) -> Scoring:
    if isinstance(sc, Scoring):
        return sc
    return Scoring(*sc)


def _band(band) -> int:
    return UMAX if band is None else int(band)


# ---- Myers ----

def distance(p, t, mode, k=NONE):
    """align.rs:87-90 `distance`; None for Option::None."""
    p, t = _b(p), _b(t)
    d = lib().sfo_distance(p, len(p), t, len(t), mode, k)
    return None if d == NONE else d


def infix_distance(p, t):
    """align.rs:655-658."""
    return distance(p, t, INFIX, NONE)


def batch_distances(p, texts, mode, k=NONE):
    """align.rs:170-200 with the 8-lane union band of batch_lanes (align.rs:552-653)."""
    p = _b(p)
    ts = [_b(t) for t in texts]
    n = len(ts)
    arr = (C.c_char_p * max(n, 1))(*ts)
    lens = (C.c_uint64 * max(n, 1))(*[len(t) for t in ts])
    out = (C.c_uint32 * max(n, 1))()
    lib().sfo_batch_distances(p, len(p), arr, lens, n, mode, k, out)
    return [None if out[i] == NONE else out[i] for i in range(n)]


def _texts(texts):
    ts = [_b(t) for t in texts]
    n = len(ts)
    return ts, n, (C.c_char_p * max(n, 1))(*ts), (C.c_uint64 * max(n, 1))(*[len(t) for t in ts])


def banded(read, texts, mode, cap, distinct=False):
    """crates/map/src/lib.rs:851-911 `Mapper::banded`: one u32 per text."""
    r = _b(read)
    ts, n, arr, lens = _texts(texts)
    out = (C.c_uint32 * max(n, 1))()
    lib().sfo_banded(r, len(r), arr, lens, n, mode, cap, int(distinct), out)
    return [out[i] for i in range(n)]


def sequential(read, texts, mode, cap):
    """crates/map/src/lib.rs:3016-3034 `sequential` (the definition `banded` is tested against)."""
    r = _b(read)
    ts, n, arr, lens = _texts(texts)
    out = (C.c_uint32 * max(n, 1))()
    lib().sfo_sequential(r, len(r), arr, lens, n, mode, cap, out)
    return [out[i] for i in range(n)]


def locate(p, t, mode, k=NONE):
    """Own definition: (distance, end, start); Nones when over k."""
    p, t = _b(p), _b(t)
    d, e, s = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().sfo_locate(p, len(p), t, len(t), mode, k, C.byref(d), C.byref(e), C.byref(s))
    f = lambda x: None if x.value == NONE else x.value
    return f(d), f(e), f(s)


# ---- Affine ----

def affine_dp(p, t, mode, sc=ONT, band=None, geom=None, clip=None, cut=None) -> Dp:
    """align.rs:2183-2208 `affine_dp` / run_core; clip = (xdrop, cell) or None."""
    p, t = _b(p), _b(t)
    out = Dp()
    has, xd, cell = (0, 0, 0) if clip is None else (1, clip[0], int(clip[1]))
    lib().sfo_affine_dp(p, len(p), t, len(t), mode, C.byref(_sc(sc)), _band(band),
                        len(t) if geom is None else geom, has, xd, cell, UMAX if cut is None else cut, C.byref(out))
    return out


def affine(p, t, mode, sc=ONT, band=None) -> int:
    p, t = _b(p), _b(t)
    return lib().sfo_affine(p, len(p), t, len(t), mode, C.byref(_sc(sc)), _band(band))


def affine_extend(p, t, sc=ONT, band=None, end_bonus=0, xdrop=2**31 - 1) -> int:
    p, t = _b(p), _b(t)
    return lib().sfo_affine_extend(p, len(p), t, len(t), C.byref(_sc(sc)), _band(band), end_bonus, xdrop)


def affine_extend_window(p, t, window, sc=ONT, band=None, end_bonus=0, xdrop=2**31 - 1):
    p, t = _b(p), _b(t)
    out = (C.c_int64 * 4)()
    lib().sfo_affine_extend_window(p, len(p), t, len(t), window, C.byref(_sc(sc)), _band(band), end_bonus, xdrop, out)
    return tuple(out)


def affine_extend_cut_bonus(p, t, sc, band, end_bonus, xdrop, diagonal, window, cut, track_cut=True,
                            bonus_within_cut=False):
    p, t = _b(p), _b(t)
    out = (C.c_int64 * 3)()
    lib().sfo_affine_extend_cut_bonus(p, len(p), t, len(t), C.byref(_sc(sc)), _band(band), end_bonus, xdrop,
                                      int(diagonal), window, UMAX if cut is None else cut, int(track_cut),
                                      int(bonus_within_cut), out)
    return tuple(out)


def local_band_ends(read, t, sc, dlo, dhi, bonus):
    r, t = _b(read), _b(t)
    return lib().sfo_local_band_ends(r, len(r), t, len(t), C.byref(_sc(sc)), dlo, dhi, bonus[0], bonus[1])


def local_band(read, t, sc, dlo, dhi, end_bonus):
    """align.rs:735-737."""
    return local_band_ends(read, t, sc, dlo, dhi, (end_bonus, end_bonus))


def align(p, t, mode, sc=ONT, band=None):
    """Own traceback definition: (score, score_col, tstart, cigar list) or None if the pass stopped."""
    p, t = _b(p), _b(t)
    cap = len(p) + len(t) + 2
    buf = (C.c_uint32 * cap)()
    s, col, ts = C.c_int32(), C.c_uint64(), C.c_uint64()
    r = lib().sfo_align(p, len(p), t, len(t), mode, C.byref(_sc(sc)), _band(band), buf, cap, C.byref(s),
                        C.byref(col), C.byref(ts))
    if r == -2:
        return None
    assert r >= 0
    return s.value, col.value, ts.value, [buf[i] for i in range(r)]


def naive_edit(p, t, mode):
    p, t = _b(p), _b(t)
    return lib().sfo_naive_edit(p, len(p), t, len(t), mode)


def naive_edit_lastrow(p, t, mode):
    p, t = _b(p), _b(t)
    out = (C.c_uint32 * (len(t) + 1))()
    lib().sfo_naive_edit_lastrow(p, len(p), t, len(t), mode, out)
    return list(out)


def naive_affine(p, t, mode, sc=ONT):
    p, t = _b(p), _b(t)
    return lib().sfo_naive_affine(p, len(p), t, len(t), mode, C.byref(_sc(sc)))


def naive_extend(p, t, sc=ONT):
    p, t = _b(p), _b(t)
    return lib().sfo_naive_extend(p, len(p), t, len(t), C.byref(_sc(sc)))


def naive_local(p, t, sc, bonus):
    p, t = _b(p), _b(t)
    return lib().sfo_naive_local(p, len(p), t, len(t), C.byref(_sc(sc)), bonus[0], bonus[1])


def naive_local_band(p, t, sc, dlo, dhi, bonus):
    p, t = _b(p), _b(t)
    return lib().sfo_naive_local_band(p, len(p), t, len(t), C.byref(_sc(sc)), dlo, dhi, bonus[0], bonus[1])


def affine_cells(m, n, mode, band=None) -> int:
    return lib().sfo_affine_cells(m, n, mode, _band(band))


# ---- batch drivers (CPU baseline) ----

def batch_myers(pat, poff, plen, txt, toff, tlen, mode, k=NONE, threads=1):
    n = len(plen)
    out = np.empty(n, np.uint32)
    a = [np.ascontiguousarray(x) for x in (pat, poff, plen, txt, toff, tlen)]
    lib().sfo_batch_myers(*[x.ctypes.data for x in a], n, mode, k, out.ctypes.data, threads)
    return out


def avx512_available() -> bool:
    return bool(lib().sfo_avx512_available())


def batch_affine(pat, poff, plen, txt, toff, tlen, mode, sc, band, threads=1, avx512=False):
    """run_core over many pairs on `threads` host threads; avx512: with the AVX-512 rows (the
    host must have AVX-512 F/BW/VL, see avx512_available)."""
    n = len(plen)
    score = np.empty(n, np.int32)
    col = np.empty(n, np.uint64)
    band = np.ascontiguousarray(np.broadcast_to(np.asarray(band, np.uint64), (n,)))
    a = [np.ascontiguousarray(x) for x in (pat, poff, plen, txt, toff, tlen)]
    args = [*[x.ctypes.data for x in a], n, mode, C.byref(_sc(sc)), band.ctypes.data, score.ctypes.data,
            col.ctypes.data, threads]
    if avx512:
        if lib().sfo_batch_affine_avx512(*args) != 0:
            raise RuntimeError("this host has no AVX-512 F/BW/VL")
    else:
        lib().sfo_batch_affine(*args)
    return score, col


# ---- CIGAR helpers ----

OPS = {0: "M", 1: "I", 2: "D", 7: "=", 8: "X"}


def cigar_str(cig) -> str:
    return "".join(f"{c >> 4}{OPS[c & 15]}" for c in cig)


# ---- the index's packed store and the tie-break's segment call ----

def pack_u64(codes) -> np.ndarray:
    """crates/index/src/lib.rs packing: base i at bits 2 (i % 32) of u64 word i / 32."""
    c = np.asarray(bytearray(_b(codes)), np.uint64) & 3
    w = np.zeros((len(c) + 31) // 32, np.uint64)
    for k in range(32):
        part = c[k::32]
        w[: len(part)] |= part << np.uint64(2 * k)
    return w


def decode_u64(words, lo, hi) -> bytes:
    """crates/index/src/lib.rs:64-90 `decode_into`: bases [lo, hi) as one code a byte."""
    return bytes(int(words[i // 32]) >> (2 * (i % 32)) & 3 for i in range(lo, hi))


def word_at(words, i) -> int:
    """crates/index/src/lib.rs:92-100 `word_at`: 32 bases from base i, words past the end read as 0."""
    q, s = i // 32, 2 * (i % 32)
    get = lambda k: int(words[k]) if k < len(words) else 0
    lo = get(q) >> s
    return lo if s == 0 else (lo | (get(q + 1) << (64 - s))) & (2**64 - 1)


def packed_eq(a, alo, b, blo, length) -> bool:
    """crates/index/src/lib.rs:104-114 `packed_eq`."""
    k = 0
    while k + 32 <= length:
        if word_at(a, alo + k) != word_at(b, blo + k):
            return False
        k += 32
    r = length - k
    return r == 0 or ((word_at(a, alo + k) ^ word_at(b, blo + k)) & ((1 << (2 * r)) - 1)) == 0


SEG_REVERSE, SEG_EXTEND, SEG_DIAGONAL, SEG_TRACK_CUT, SEG_BONUS_WITHIN_CUT = 1, 2, 4, 8, 16


def tiebreak_segment(words, read, cands, mode, flags, window, cut, sc, band, end_bonus, xdrop):
    """crates/map/src/lib.rs:1307-1492 `seg` with `tb.affine` set and `affine_within` unset, as
    per-candidate results (the reference's deltas are their negations): cands = [(store offset,
    length)]; returns ([(whole, cut, text bases)], rep) where rep[c] is the first candidate with an
    equal text (lib.rs:1337-1345, by `sequence_eq`, lib.rs:135-137).  Distinct texts are decoded,
    reversed with SEG_REVERSE together with the read (lib.rs:1347-1365), and scored one at a time:
    align.rs:903-949 for SEG_EXTEND with the window the larger of `window` and the longest text
    (lib.rs:1446-1463), else align.rs:854-861 as (x, x, 0) (lib.rs:1464-1469).  Scoring sorted
    texts together equals scoring them one at a time (align.rs:2210-2213)."""
    rev = bool(flags & SEG_REVERSE)
    rep, distinct = [], []
    for c, (off, n) in enumerate(cands):
        r = next((d for d in distinct if cands[d][1] == n and packed_eq(words, cands[d][0], words, off, n)), c)
        if r == c:
            distinct.append(c)
        rep.append(r)
    r_ = _b(read)[::-1] if rev else _b(read)
    texts = {d: (decode_u64(words, cands[d][0], cands[d][0] + cands[d][1])[::-1] if rev
                 else decode_u64(words, cands[d][0], cands[d][0] + cands[d][1])) for d in distinct}
    res = {}
    if flags & SEG_EXTEND:
        w = max([window] + [len(t) for t in texts.values()])
        for d, t in texts.items():
            res[d] = affine_extend_cut_bonus(r_, t, sc, band, end_bonus, xdrop, bool(flags & SEG_DIAGONAL), w,
                                             None if cut == UMAX else cut, bool(flags & SEG_TRACK_CUT),
                                             bool(flags & SEG_BONUS_WITHIN_CUT))
    else:
        for d, t in texts.items():
            x = affine(r_, t, mode, sc, band)
            res[d] = (x, x, 0)
    return [tuple(res[r]) for r in rep], rep


# The C form of `tiebreak_segment` over many segments (the CPU baseline of the tie-break call mix).
TB_SEGMENT = np.dtype([("read_off", np.uint64), ("read_len", np.uint32), ("cand0", np.uint32), ("ncand", np.uint32),
                       ("mode", np.int32), ("flags", np.uint32), ("window", np.uint64), ("cut", np.uint64)],
                      align=True)


def tiebreak_batch(words, bases, reads, segs, cand_off, cand_len, sc, band, end_bonus, xdrop, threads=1,
                   avx512=False):
    """sfo_tiebreak_segments: segs a TB_SEGMENT array; returns (out (n, 3) int64, rep uint32)."""
    w = np.ascontiguousarray(words, np.uint64)
    r = np.frombuffer(_b(reads), np.uint8) if len(reads) else np.zeros(1, np.uint8)
    sg = np.ascontiguousarray(segs, TB_SEGMENT)
    co = np.ascontiguousarray(cand_off, np.uint64)
    cl = np.ascontiguousarray(cand_len, np.uint32)
    n = len(co)
    out = np.zeros((max(n, 1), 3), np.int64)
    rep = np.zeros(max(n, 1), np.uint32)
    st = lib().sfo_tiebreak_segments(w.ctypes.data, bases, r.ctypes.data, sg.ctypes.data, len(sg), co.ctypes.data,
                                     cl.ctypes.data, C.byref(_sc(sc)), _band(band), end_bonus, xdrop,
                                     out.ctypes.data, rep.ctypes.data, threads, int(avx512))
    if st != 0:
        raise RuntimeError("this host has no AVX-512 F/BW/VL")
    return out[:n], rep[:n]
