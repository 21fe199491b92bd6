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
from spearfish_amd.align import Mode
from tb_cases import TBS, _seq, make_segments, make_store, oracle

pytestmark = pytest.mark.gpu




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


def test_batch_api_equals_one_shot_and_reruns():
    # sf_tb_batch_*: made once, run with two parameter sets in turn (the second twice); each run's
    # results equal the one-shot call's with the same parameters, and the batch outlives its store.
    rng = random.Random(7)
    codes, genes = make_store(rng, genes=20)
    store = T.PackedStore.from_codes(codes)
    tb0, tb1 = TBS["default"], TBS["hifi_band64"]
    reads, segs = make_segments(rng, codes, genes, tb0)
    arr = np.zeros(len(segs), T.SEGMENT)
    off, ln = [], []
    for i, s in enumerate(segs):
        arr[i] = (s.read_off, s.read_len, len(off), len(s.cands), int(s.mode), T.flags(s.mode, s.rev, tb0),
                  tb0.window(s.read_len), 2**64 - 1 if s.cut is None else s.cut)
        off += [c[0] for c in s.cands]
        ln += [c[1] for c in s.cands]
    want0 = T.segments_arrays(store, reads, arr, off, ln, tb0)
    want1 = T.segments_arrays(store, reads, arr, off, ln, tb1)
    b = T.TbBatch(store, reads, arr, off, ln)
    store.close()
    for tb, want in ((tb0, want0), (tb1, want1), (tb1, want1), (tb0, want0)):
        b.run(tb)
        out, rep = b.results()
        assert out.tolist() == want[0].tolist() and rep.tolist() == want[1].tolist()
    b.close()
