"""`sf_tiebreak_segments` and `sf_store_codes_eq` on the device against the oracle's restatement of
the tie-break's `seg` (crates/map/src/lib.rs:1307-1492) and of the index's packed equality
(crates/index/src/lib.rs:104-114, 135-137), on a synthetic isoform store: genes of exons,
transcripts that keep some of them, packed back to back in the index's u64 words.  The segments are
the reference's three kinds: right read ends (Prefix, windows of `tb.window(read length)` bases,
cut at the shortest window that stops at its transcript's end), left read ends (the same, reversed)
and interior stretches (Global), plus Infix; tie-break settings vary per call."""
import ctypes
import random

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import _native as N
from spearfish_amd import tiebreak as T
from spearfish_amd.align import Affine, Mode

pytestmark = pytest.mark.gpu


def _seq(rng, n):
    return bytes(rng.randrange(4) for _ in range(n))


def _mutate(rng, s, rate=0.08):
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < rate / 3:  # deletion
            continue
        if r < 2 * rate / 3:  # insertion before the base
            out.append(rng.randrange(4))
        out.append(rng.randrange(4) if r >= 2 * rate / 3 and r < rate else b)  # substitution
    return bytes(out)


def make_store(rng, genes=30):
    codes = bytearray()
    out = []
    for _ in range(genes):
        exons = [_seq(rng, rng.randrange(6, 140)) for _ in range(rng.randrange(2, 7))]
        isos = []
        for _ in range(rng.randrange(2, 8)):
            keep = [e for e in range(len(exons)) if rng.random() < 0.7] or [rng.randrange(len(exons))]
            codes += _seq(rng, rng.randrange(0, 40))  # intergenic bases: offsets at any word position
            off, pos, p = len(codes), {}, 0
            for e in keep:
                pos[e] = p
                p += len(exons[e])
                codes += exons[e]
            isos.append((off, p, pos))
        out.append((exons, isos))
    return bytes(codes), out


def make_segments(rng, codes, genes, tb):
    """Segments of the three kinds; returns (reads buffer, [Segment])."""
    reads, segs = bytearray(), []

    def add_read(r):
        reads.extend(r)
        return len(reads) - len(r), len(r)

    for exons, isos in genes:
        for e in range(len(exons)):
            has = [iso for iso in isos if e in iso[2]]
            if not has:
                continue
            kind = rng.randrange(4)
            if kind == 0:  # right end from exon e's start
                base = codes[has[0][0] + has[0][2][e]: has[0][0] + has[0][1]]
                read = _mutate(rng, base[: rng.randrange(0, 90)])
                w = tb.window(len(read))
                cands = [(off + pos[e], min(w, ln - pos[e])) for off, ln, pos in has]
                short = [ln - pos[e] < w for off, ln, pos in has]
                lens = [c[1] for c in cands]
                cut = min((l for l, s in zip(lens, short) if s), default=None)
                if cut is None or not (tb.flank_min > 0 or tb.end_bonus_cut) or not any(l > cut for l in lens):
                    cut = None
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Prefix, False, cut))
            elif kind == 1:  # left end up to exon e's end, aligned reversed
                ends = [(off, pos[e] + len(exons[e])) for off, ln, pos in has]
                base = codes[ends[0][0]: ends[0][0] + ends[0][1]]
                read = _mutate(rng, base[max(0, len(base) - rng.randrange(0, 90)):])
                w = tb.window(len(read))
                cands = [(off + max(0, end - w), min(w, end)) for off, end in ends]
                lens = [c[1] for c in cands]
                short = [end < w for off, end in ends]
                cut = min((l for l, s in zip(lens, short) if s), default=None)
                if cut is None or not (tb.flank_min > 0 or tb.end_bonus_cut) or not any(l > cut for l in lens):
                    cut = None
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Prefix, True, cut))
            else:  # an interior stretch from exon e's start, some running into the next exon
                extra = rng.randrange(0, 12)
                cands = [(off + pos[e], min(len(exons[e]) + extra, ln - pos[e])) for off, ln, pos in has]
                read = _mutate(rng, exons[e])
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Global if kind == 2 else Mode.Infix))
    return bytes(reads), segs


def flags_of(s, tb):
    f = O.SEG_REVERSE if s.rev else 0
    if s.mode == Mode.Prefix and tb.clip_ends:
        f |= O.SEG_EXTEND
    f |= (O.SEG_DIAGONAL if tb.end_diagonal else 0) | (O.SEG_TRACK_CUT if tb.flank_min > 0 else 0)
    return f | (O.SEG_BONUS_WITHIN_CUT if tb.end_bonus_cut else 0)


def oracle(words, reads, segs, tb):
    res, reps = [], []
    for s in segs:
        r, rep = O.tiebreak_segment(words, reads[s.read_off: s.read_off + s.read_len], list(s.cands), int(s.mode),
                                    flags_of(s, tb), tb.window(s.read_len), O.UMAX if s.cut is None else s.cut,
                                    (tb.affine.mat, tb.affine.mis, tb.affine.open, tb.affine.ext),
                                    tb.band, tb.end_bonus, tb.xdrop)
        res += [tuple(x) for x in r]
        reps += rep
    return res, reps


TBS = {
    "default": T.TieBreak(),
    "hifi_band64": T.TieBreak(affine=Affine.HIFI, band=64),
    "diagonal": T.TieBreak(end_diagonal=True),
    "flank_min": T.TieBreak(flank_min=12, end_bonus_cut=False),
    "flank_min_bonus_cut": T.TieBreak(flank_min=12),
    "no_clip_ends": T.TieBreak(clip_ends=False, band=8),
    "xdrop_small_band_wide": T.TieBreak(xdrop=8, band=300, end_bonus=9),
}


@pytest.mark.parametrize("name", list(TBS))
def test_segments_equal_oracle(name):
    tb = TBS[name]
    rng = random.Random(100 + list(TBS).index(name))
    codes, genes = make_store(rng)
    words = O.pack_u64(codes)  # the index's own u64 words
    reads, segs = make_segments(rng, codes, genes, tb)
    store = T.PackedStore(words, len(codes))
    whole, cut, cols, rep = T.segments(store, reads, segs, tb)
    want, want_rep = oracle(words, reads, segs, tb)
    got = list(zip(whole.tolist(), cut.tolist(), cols.tolist()))
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, (name, len(bad), len(want), [(got[i], want[i]) for i in bad[:5]])
    assert rep.tolist() == want_rep
    assert len(set(rep.tolist())) < len(rep)  # some texts repeat within segments


def test_one_call_equals_calls_per_segment():
    tb = T.TieBreak()
    rng = random.Random(3)
    codes, genes = make_store(rng, genes=12)
    store = T.PackedStore.from_codes(codes)
    reads, segs = make_segments(rng, codes, genes, tb)
    allr = T.segments(store, reads, segs, tb)
    k = 0
    for s in segs:
        one = T.segments(store, reads, [s], tb)
        for a, b in zip(allr, one):
            assert a[k:k + len(s.cands)].tolist() == b.tolist()
        k += len(s.cands)


def test_sequence_eq_on_device():
    rng = random.Random(4)
    codes = bytearray(_seq(rng, 5000))
    for _ in range(40):  # repeats at unrelated word alignments
        a, n = rng.randrange(0, 4000), rng.randrange(0, 300)
        b = rng.randrange(0, 5000 - n)
        codes[b:b + n] = codes[a:a + n]
    codes = bytes(codes)
    words = O.pack_u64(codes)
    store = T.PackedStore(words, len(codes))
    A, B, L = [], [], []
    for _ in range(3000):
        n = rng.choice([0, 1, 31, 32, 33, 64, rng.randrange(0, 400)])
        a = rng.randrange(0, len(codes) - n + 1)
        b = a if rng.random() < 0.2 else rng.randrange(0, len(codes) - n + 1)
        A.append(a), B.append(b), L.append(n)
    # Equal ranges: each repeat planted above, found by a search over the decoded bases.
    for _ in range(200):
        n = rng.randrange(1, 100)
        a = rng.randrange(0, len(codes) - n)
        b = codes.find(codes[a:a + n], a + 1)
        if b >= 0:
            A.append(a), B.append(b), L.append(n)
    got = store.sequence_eq(A, B, L)
    want = [codes[a:a + n] == codes[b:b + n] for a, b, n in zip(A, B, L)]
    assert got.tolist() == want
    assert [O.packed_eq(words, a, words, b, n) for a, b, n in zip(A[:500], B[:500], L[:500])] == want[:500]
    assert sum(want) > 300


def test_empty_reads_and_texts():
    # An empty read against texts (the reference's insertion-only Global shortcut, lib.rs:1321-1333:
    # open + ext * length, 0 for an empty text), empty texts in every kind, and no segments at all.
    tb = T.TieBreak()
    rng = random.Random(6)
    codes = _seq(rng, 400)
    store = T.PackedStore.from_codes(codes)
    cands = [(10, 0), (20, 5), (30, 17), (40, 0)]
    whole, cut, cols, rep = T.segments(store, b"", [T.Segment(0, 0, cands, Mode.Global)], tb)
    sc = tb.affine
    assert (-whole).tolist() == [0 if n == 0 else sc.open + sc.ext * n for _, n in cands]
    reads = _seq(rng, 30)
    segs = [T.Segment(0, 30, cands, m, rev, cut) for m in (Mode.Infix, Mode.Prefix, Mode.Global)
            for rev in (False, True) for cut in (None, 3)]
    segs += [T.Segment(5, 0, [(0, 0), (100, 50)], Mode.Prefix, True, None)]
    got = T.segments(store, reads, segs, tb)
    want, want_rep = oracle(O.pack_u64(codes), reads, segs, tb)
    assert list(zip(*[x.tolist() for x in got[:3]])) == want
    assert got[3].tolist() == want_rep
    assert [x.tolist() for x in T.segments(store, reads, [], tb)] == [[], [], [], []]


def test_invalid_arguments_are_refused():
    tb = T.TieBreak()
    store = T.PackedStore.from_codes(bytes(100))
    bad = [
        [T.Segment(0, 4, [(90, 20)], Mode.Global)],  # candidate past the store's end
        [T.Segment(0, 40, [(0, 10)], Mode.Global)],  # read slice past the reads
    ]
    for segs in bad:
        with pytest.raises(N.NativeError) as e:
            T.segments(store, bytes(8), segs, tb)
        assert e.value.status == N.SF_EINVAL
    with pytest.raises(N.NativeError):  # an extension that is not Prefix
        lib = N.load()
        arr = (N.TbSegment * 1)(N.TbSegment(0, 4, 0, 1, int(Mode.Global), N.SEG_EXTEND, 30, 2**64 - 1))
        co, cl = np.zeros(1, np.uint64), np.full(1, 5, np.uint32)
        out, rep = np.zeros(3, np.int32), np.zeros(1, np.uint32)
        p = N.TbParams(tb.affine.c(), 16, 5, 50)
        N.check(lib.sf_tiebreak_segments(store._h, bytes(8), 8, arr, 1, co.ctypes.data, cl.ctypes.data, 1,
                                         ctypes.byref(p), out.ctypes.data, rep.ctypes.data))
