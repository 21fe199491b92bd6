"""The oracle's packed store and tie-break segment call (oracle/sf_oracle.py `packed_eq`,
`tiebreak_segment`), pinned:

* `packed_eq` equals comparing the decoded bases, on the inputs of the index's own property test
  (crates/index/src/lib.rs:1081-1101 `packed_eq_matches_scalar`: sequences of 1..300 bases, any
  alignment, an optional flipped base);
* a segment of one candidate equals the independent full-matrix restatement (tests/fullmatrix.py)
  of align.rs:903-949 (extensions) or the textbook Gotoh score (other segments);
* equal texts get equal results and `rep` names the first equal candidate (decoded comparison);
* a reversed segment equals the forward call on the reversed read against a store holding the
  reversed texts, so both the read and the texts are reversed;
* with texts of several lengths the band is laid out for the segment's longest text
  (lib.rs:1446-1451): a planted mistake (each text's own length) fails the comparison.
"""
import random

import numpy as np

import fullmatrix as FM
from oracle import sf_oracle as O

UMAX = O.UMAX


def test_packed_eq_matches_decoded_comparison():
    rng = random.Random(5)
    for _ in range(400):
        n = rng.randrange(1, 300)
        bases = bytes(rng.randrange(4) for _ in range(n))
        other = bytearray(bases)
        if rng.random() < 0.5:
            other[rng.randrange(n)] ^= 1
        a, b = rng.randrange(n), rng.randrange(n)
        ln = min(rng.randrange(300), n - a, n - b)
        w, v = O.pack_u64(bases), O.pack_u64(bytes(other))
        assert O.packed_eq(w, a, v, b, ln) == (bases[a:a + ln] == bytes(other)[b:b + ln])
        assert O.packed_eq(w, a, w, a, ln)
        assert O.decode_u64(w, a, a + ln) == bases[a:a + ln]


def test_word_at_reads_zero_past_the_end():
    w = O.pack_u64(bytes([3] * 40))
    assert O.word_at(w, 32) == (1 << 16) - 1  # 8 bases of 3, then zeros
    assert O.word_at(w, 64) == 0


def _store(rng, texts):
    """A store holding the texts back to back, with random bases between them."""
    codes, offs = bytearray(), []
    for t in texts:
        codes += bytes(rng.randrange(4) for _ in range(rng.randrange(0, 5)))
        offs.append((len(codes), len(t)))
        codes += bytes(t)
    return O.pack_u64(bytes(codes)), offs


def _related(rng, n):
    a, b = FM_related(rng, n)
    return bytes(a), bytes(b)


def FM_related(rng, n):
    from test_oracle_p2 import related
    return related(rng, n)


SC = [O.ONT, O.HIFI, (3, 2, 5, 1)]


def test_single_candidate_equals_the_full_matrix():
    rng = random.Random(7)
    for _ in range(600):
        p, t = _related(rng, rng.randrange(0, 26))
        if rng.random() < 0.5:
            t = t + bytes(rng.randrange(4) for _ in range(rng.randrange(0, 10)))
        words, offs = _store(rng, [t])
        sc, band = rng.choice(SC), rng.choice([0, 1, 3, 8])
        bonus, xdrop = rng.choice([0, 5]), rng.choice([0, 3, 50, 2**31 - 1])
        window = rng.randrange(0, len(t) + 8)
        cut = rng.choice([UMAX, rng.randrange(0, len(t) + 2)])
        if rng.random() < 0.6:
            flags = O.SEG_EXTEND | rng.choice([0, O.SEG_DIAGONAL]) | rng.choice([0, O.SEG_TRACK_CUT]) | \
                rng.choice([0, O.SEG_BONUS_WITHIN_CUT])
            (got,), rep = O.tiebreak_segment(words, p, offs, O.PREFIX, flags, window, cut, sc, band, bonus, xdrop)
            want = FM.extend_cut_bonus(p, t, sc, band, bonus, xdrop, bool(flags & O.SEG_DIAGONAL), max(window, len(t)),
                                       None if cut == UMAX else cut, bool(flags & O.SEG_TRACK_CUT),
                                       bool(flags & O.SEG_BONUS_WITHIN_CUT))
        else:
            mode = rng.choice([O.INFIX, O.PREFIX, O.GLOBAL])
            (got,), rep = O.tiebreak_segment(words, p, offs, mode, 0, window, cut, sc, None, bonus, xdrop)
            x = O.naive_affine(p, t, mode, sc)
            want = (x, x, 0)
        assert rep == [0]
        assert tuple(got) == tuple(want)


def _segment_case(rng):
    base = [bytes(rng.randrange(4) for _ in range(rng.randrange(0, 30))) for _ in range(rng.randrange(1, 4))]
    texts = [rng.choice(base) if rng.random() < 0.6 else _related(rng, rng.randrange(0, 30))[1]
             for _ in range(rng.randrange(1, 7))]
    read = _related(rng, rng.randrange(0, 30))[0]
    return read, texts


def test_equal_texts_share_results_and_rep_names_the_first():
    rng = random.Random(8)
    for _ in range(300):
        read, texts = _segment_case(rng)
        words, offs = _store(rng, texts)
        flags = rng.choice([0, O.SEG_EXTEND | O.SEG_BONUS_WITHIN_CUT, O.SEG_EXTEND | O.SEG_TRACK_CUT])
        mode = O.PREFIX if flags else rng.choice([O.INFIX, O.GLOBAL])
        res, rep = O.tiebreak_segment(words, read, offs, mode, flags, len(read) + 20, rng.choice([UMAX, 10]), O.ONT,
                                      16, 5, 50)
        for c, t in enumerate(texts):
            assert rep[c] == next(d for d in range(c + 1) if texts[d] == t)
            assert res[c] == res[rep[c]]


def test_reversed_segment_reverses_read_and_texts():
    rng = random.Random(9)
    for _ in range(300):
        read, texts = _segment_case(rng)
        words, offs = _store(rng, texts)
        rwords, roffs = _store(random.Random(1), [t[::-1] for t in texts])
        flags = rng.choice([0, O.SEG_EXTEND, O.SEG_EXTEND | O.SEG_TRACK_CUT | O.SEG_BONUS_WITHIN_CUT])
        mode = O.PREFIX if flags or rng.random() < 0.4 else rng.choice([O.INFIX, O.GLOBAL])
        cut = rng.choice([UMAX, rng.randrange(0, 20)])
        args = (mode, flags, len(read) + 20, cut, O.ONT, rng.choice([2, 16]), 5, 50)
        got = O.tiebreak_segment(words, read, offs, args[0], args[1] | O.SEG_REVERSE, *args[2:])
        want = O.tiebreak_segment(rwords, read[::-1], roffs, *args)
        assert got == want


def test_extension_band_is_laid_out_for_the_longest_text():
    rng = random.Random(10)
    wrong = 0
    for _ in range(300):
        # A read with an insertion longer than the band against a text without it, beside a text as
        # long as the read: laid out for the longer text, the band cannot follow the insertion.
        t = bytes(rng.randrange(4) for _ in range(rng.randrange(10, 24)))
        a = rng.randrange(2, len(t) - 2)
        read = t[:a] + bytes(rng.randrange(4) for _ in range(rng.randrange(3, 9))) + t[a:]
        texts = [t, read + bytes(rng.randrange(4) for _ in range(rng.randrange(0, 4)))]
        rng.shuffle(texts)
        words, offs = _store(rng, texts)
        window = rng.randrange(0, 6)
        res, _ = O.tiebreak_segment(words, read, offs, O.PREFIX, O.SEG_EXTEND, window, UMAX, O.ONT, 2, 5, 50)
        longest = max(len(t) for t in texts)
        for c, t in enumerate(texts):
            want = FM.extend_cut_bonus(read, t, O.ONT, 2, 5, 50, False, max(window, longest), None, False, False)
            assert tuple(res[c]) == want
            # The planted mistake: the band laid out for each text's own length.
            wrong += FM.extend_cut_bonus(read, t, O.ONT, 2, 5, 50, False, max(window, len(t)), None, False,
                                         False) != want
    assert wrong > 0


def test_c_batch_equals_the_restatement():
    # oracle/sf_oracle.c `sfo_tiebreak_segments` (the CPU baseline's driver: threads over segments,
    # scalar or the reference's x86 rows) against `tiebreak_segment` above, on the isoform cases.
    from tb_cases import TBS, flags_of, make_segments, make_store, oracle

    for name, tb in TBS.items():
        rng = random.Random(200 + list(TBS).index(name))
        codes, genes = make_store(rng, genes=12)
        words = O.pack_u64(codes)
        reads, segs = make_segments(rng, codes, genes, tb)
        want, want_rep = oracle(words, reads, segs, tb)
        arr = np.zeros(len(segs), O.TB_SEGMENT)
        off, ln = [], []
        for i, s in enumerate(segs):
            arr[i] = (s.read_off, s.read_len, len(off), len(s.cands), int(s.mode), flags_of(s, tb),
                      tb.window(s.read_len), O.UMAX if s.cut is None else s.cut)
            off += [c[0] for c in s.cands]
            ln += [c[1] for c in s.cands]
        sc = (tb.affine.mat, tb.affine.mis, tb.affine.open, tb.affine.ext)
        for avx in ([False, True] if O.avx512_available() else [False]):
            out, rep = O.tiebreak_batch(words, len(codes), reads, arr, off, ln, sc, tb.band, tb.end_bonus, tb.xdrop,
                                        threads=4, avx512=avx)
            assert [tuple(x) for x in out.tolist()] == want, (name, avx)
            assert rep.tolist() == want_rep
