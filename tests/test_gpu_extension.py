"""GPU parity of the passes with per-row bookkeeping: affine_dp with X-drop clip, best cell, cut and
band layout (align.rs:2183-2208, 2446-2666), the extension wrappers (align.rs:710-729, 880-982) and
the local band (align.rs:731-852), through the C ABI, against the reference's known answers and
the oracle.
"""
import random

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A

pytestmark = pytest.mark.gpu

ONT = A.Affine.ONT
IMAX = 2**31 - 1
UMAX = A.USIZE_MAX


def opt(b):
    return None if b == UMAX else b


# ---- the reference's known answers, on the device ----

def test_local_band_examples():
    # align.rs:3016-3043
    t = [3, 3, 0, 1, 2, 3, 0, 1, 2, 2, 2]
    p = [0, 1, 2, 3, 0, 1]
    assert A.local_band(p, t, ONT, 2, 2, 5) == 12 + 10
    assert A.local_band(p, t, ONT, -6, 11, 5) == 12 + 10
    assert A.local_band(p, t, ONT, 0, 0, 0) == 0
    assert A.local_band(p, t, ONT, 6, 6, 0) == 6
    junk = [2, 2, 2, 0, 1, 2, 3, 0, 1, 3, 3, 3]
    assert A.local_band(junk, t, ONT, -12, 11, 5) == 12
    assert A.local_band([4, 4, 4], [0, 0, 0], ONT, -3, 3, 0) == 0
    assert A.local_band([], t, ONT, -3, 3, 5) == 0
    assert A.local_band(p, t, ONT, 3, 2, 5) == 0
    assert A.local_band_ends(p, t, ONT, 2, 2, (5, 0)) == 12 + 5
    assert A.local_band_ends(p, t, ONT, 2, 2, (0, 3)) == 12 + 3
    assert A.local_band_ends([0, 1, 2, 2], [0, 1, 2], ONT, 0, 0, (0, 5)) == 6
    assert A.local_band_ends([0, 1, 2, 2], [0, 1, 2], ONT, 0, 0, (5, 0)) == 6 + 5


def test_affine_extend_examples():
    # align.rs:3118-3129
    t = [0, 1, 2, 3, 0, 1, 2, 3]
    p = [0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3]
    assert A.affine_extend(p, t, ONT, UMAX, 0, IMAX) == O.naive_extend(p, t, O.ONT)
    assert A.affine_extend([], t, ONT, UMAX, 0, IMAX) == 0


def test_affine_extend_end_examples():
    # align.rs:3131-3145
    def end(p, t, bonus):
        x, i, j, _ = A.affine_extend_window(p, t, len(t), ONT, 16, bonus, 50)
        return (x, i, j)

    t = [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
    assert end([0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3], t, 0) == (12, 6, 6)
    assert end([0, 1, 2, 0, 1, 2, 3], t, 5) == (14 - 6 + 5, 7, 8)
    assert end([3, 3, 3], [0, 0, 0], 0) == (0, 0, 0)


def test_affine_extend_window_cuts():
    # align.rs:3147-3166
    t = [((i * 5 + i // 4) % 4) for i in range(140)]
    p = t[:70] + [[3, 3, 0, 0][i % 4] for i in range(40)]
    r = A.affine_extend_window(p, t, len(t), ONT, 16, 5, 30)
    assert r == O.affine_extend_window(p, t, len(t), O.ONT, 16, 5, 30)
    assert r[3] < len(t), r
    for cut in (r[3], r[3] + 1, len(t) - 1):
        assert A.affine_extend_window(p, t[:cut], len(t), ONT, 16, 5, 30)[0] == r[0]
    short = A.affine_extend_window(p, t[:40], len(t), ONT, 16, 5, 30)
    assert (short[2], short[3]) == (40, 40)


def test_affine_extend_end_reads_a_prefix():
    # align.rs:3168-3183
    t = [((i * 7 + i // 3) % 4) for i in range(120)]
    p = t[:60] + [[3, 3, 0, 0][i % 4] for i in range(40)]
    r = A.affine_extend_window(p, t, len(t), ONT, 16, 5, 30)
    assert r[3] < len(t), r
    u = list(t)
    for q in range(r[3], len(u)):
        u[q] = 3 - u[q]
    assert A.affine_extend_window(p, u, len(u), ONT, 16, 5, 30) == r


# ---- random inputs against the oracle ----

def related(rng, n, err=0.1):
    t = [rng.randrange(4) for _ in range(n)]
    p = []
    for x in t:
        u = rng.random()
        if u < err / 2:
            p.append((x + 1 + rng.randrange(3)) & 3)
        elif u < 3 * err / 4:
            p += [x, rng.randrange(4)]
        elif u < err:
            pass
        else:
            p.append(x)
    return p, t


def extension_pairs(rng, count, maxlen=300):
    """Extension-shaped pairs: a read end that aligns for a while, then turns to junk."""
    P, T = [], []
    for _ in range(count):
        n = rng.randrange(0, maxlen)
        p, t = related(rng, n, rng.choice([0.02, 0.08, 0.15]))
        if rng.random() < 0.5:
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(0, 80))]
        if rng.random() < 0.2:
            t = t[: rng.randrange(len(t) + 1)]
        P.append(p)
        T.append(t)
    return P, T


@pytest.mark.parametrize("mode", [A.Mode.Prefix, A.Mode.Global, A.Mode.Infix])
def test_affine_dp_clip_cut_geom(mode):
    rng = random.Random(71 + int(mode))
    maxlen = 120 if mode == A.Mode.Infix else 300
    P, T = extension_pairs(rng, 240, maxlen)
    n = len(P)
    bands = np.array([rng.choice([0, 1, 4, 16, 40, UMAX]) for _ in range(n)], np.uint64)
    geoms = np.array([max(len(T[i]), rng.choice([0, len(T[i]), len(T[i]) + 30, 400])) if rng.random() < 0.6
                      else min(len(P[i]), len(T[i])) for i in range(n)], np.uint64)
    cuts = np.array([rng.choice([0, 5, len(T[i]) // 2, len(T[i]), len(T[i]) + 7, 2**40, UMAX]) for i in range(n)],
                    np.uint64)
    ps = A.PairSet.from_seqs(P, T)
    for clip in (None, (0, True), (30, True), (50, False), (IMAX, True)):
        for use_geom in (False, True):
            for use_cut in (False, True):
                got = A.affine_dp_batch(ps, mode, ONT, bands, geoms if use_geom else None, clip,
                                        cuts if use_cut else None)
                for i in range(n):
                    want = O.affine_dp(P[i], T[i], int(mode), O.ONT, opt(int(bands[i])),
                                       int(geoms[i]) if use_geom else None, clip,
                                       int(cuts[i]) if use_cut else None)
                    w = (want.score, want.score_col, want.clip_best, want.clip_i, want.clip_j, want.text_read,
                         want.last_row, want.cut_best, want.cut_score)
                    assert tuple(got[i]) == w, (i, len(P[i]), len(T[i]), int(bands[i]), clip, use_geom, use_cut)


def test_affine_dp_other_scorings_and_lanes():
    # HIFI and an open = 0 scoring, with bands wide enough for the 64-lane shape too.
    rng = random.Random(83)
    P, T = extension_pairs(rng, 120, 700)
    n = len(P)
    bands = np.array([rng.choice([8, 100, 300]) for _ in range(n)], np.uint64)
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.HIFI, A.Affine(1, 1, 0, 1), A.Affine(3, 2, 5, 1)):
        got = A.affine_dp_batch(ps, A.Mode.Prefix, sc, bands, None, (40, True), None)
        for i in range(n):
            want = O.affine_dp(P[i], T[i], O.PREFIX, (sc.mat, sc.mis, sc.open, sc.ext), int(bands[i]), None,
                               (40, True), None)
            assert (got[i].score, got[i].clip_best, got[i].clip_i, got[i].clip_j, got[i].text_read,
                    got[i].last_row) == (want.score, want.clip_best, want.clip_i, want.clip_j, want.text_read,
                                         want.last_row), (i, sc)


def test_extension_wrappers():
    rng = random.Random(89)
    for trial in range(12):
        n = rng.randrange(40, 200)
        core, _ = related(rng, n, 0.05)
        pattern = core + [rng.randrange(4) for _ in range(rng.randrange(0, 30))]
        texts = []
        base = [rng.randrange(4) for _ in range(n + 60)]
        base[:len(core)] = core[:len(base)]
        for _ in range(rng.randrange(1, 8)):
            texts.append(base[: rng.randrange(0, len(base) + 1)])
        texts.sort()
        band, bonus, xdrop = rng.choice([8, 16, 32]), rng.choice([0, 5]), rng.choice([20, 50, IMAX])
        window = max(len(t) for t in texts) + rng.choice([0, 10])
        cut = rng.choice([UMAX, 0, n // 2, n])
        for diagonal in (False, True):
            for track_cut in (False, True):
                for within in (False, True):
                    got = A.affine_extend_sorted_cut_bonus(pattern, texts, ONT, band, bonus, xdrop, diagonal, window,
                                                           cut, track_cut, within)
                    want = [O.affine_extend_cut_bonus(pattern, t, O.ONT, band, bonus, xdrop, diagonal, window,
                                                      None if cut == UMAX else cut, track_cut, within)
                            for t in texts]
                    assert [tuple(g) for g in got] == [tuple(w) for w in want], (trial, diagonal, track_cut, within)
        for t in texts:
            assert A.affine_extend(pattern, t, ONT, band, bonus, xdrop) == O.affine_extend(pattern, t, O.ONT, band,
                                                                                          bonus, xdrop)
            assert A.affine_extend_window(pattern, t, window, ONT, band, bonus, xdrop) == \
                O.affine_extend_window(pattern, t, window, O.ONT, band, bonus, xdrop)


def test_local_band_random():
    rng = random.Random(97)
    P, T, LO, HI, want = [], [], [], [], []
    for _ in range(300):
        n = rng.randrange(0, 400)
        t = [rng.randrange(4) for _ in range(n)]
        m = rng.randrange(0, 300)
        start = rng.randrange(-50, n + 50)
        read = [t[start + k] if 0 <= start + k < n and rng.random() > 0.1 else rng.randrange(4) for k in range(m)]
        for k in range(m):
            if rng.random() < 0.03:
                read[k] = 4 + rng.randrange(3)  # N and other codes: match nothing
        c = start
        w = rng.choice([0, 1, 5, 40, 150])
        lo, hi = c - w, c + w + rng.randrange(0, 3)
        if rng.random() < 0.1:
            lo, hi = hi, lo  # an empty range
        if rng.random() < 0.1:
            lo, hi = n + 5, n + 20  # past the text
        P.append(read)
        T.append(t)
        LO.append(lo)
        HI.append(hi)
    for b0, b1 in ((0, 0), (5, 5), (7, 0), (0, 9)):
        ps = A.PairSet.from_seqs(P, T)
        got = A.local_band_ends_batch(ps, ONT, np.array(LO, np.int64), np.array(HI, np.int64), (b0, b1))
        for i in range(len(P)):
            w = O.local_band_ends(P[i], T[i], O.ONT, LO[i], HI[i], (b0, b1))
            assert int(got[i]) == w, (i, len(P[i]), len(T[i]), LO[i], HI[i], b0, b1)


@pytest.fixture
def pass_policy():
    """Passes in the 16-bit packed layout where its range holds, on 8 lanes where the band fits
    256 slots (1, the default) or on 16 lanes (2), or in 32 bits (0)."""
    lib = A.N.load()

    def use(policy):
        A.N.check(lib.sf_set_affine_pass_packed(int(policy)))

    yield use
    A.N.check(lib.sf_set_affine_pass_packed(1))


@pytest.mark.parametrize("mode", [A.Mode.Prefix, A.Mode.Global, A.Mode.Infix])
def test_affine_dp_packed_pass_equals_32bit_and_oracle(mode, pass_policy):
    # The packed pass (affine_pk_pass.hip, 8 lanes a pair where the band fits 256 slots, and 16
    # lanes) against the 32-bit pass on the same inputs, every Dp field: bands from 0 to the
    # packed pass's 512 slots (every class, 2 to 32 slots a lane),
    # band layouts for other text lengths, cuts inside, at and past the text, X-drops that stop
    # early or never, the best cell tracked or not; and against the oracle on a sample.
    rng = random.Random(131 + int(mode))
    maxlen = 150 if mode == A.Mode.Infix else 500
    P, T = extension_pairs(rng, 360, maxlen)
    n = len(P)
    bands = np.array([rng.choice([0, 1, 7, 15, 16, 17, 31, 40, 63, 100, 127, 200, 255]) for _ in range(n)], np.uint64)
    geoms = np.array([max(0, len(T[i]) + rng.choice([-40, -3, 0, 0, 9, 120])) for i in range(n)], np.uint64)
    cuts = np.array([rng.choice([0, 3, len(T[i]) // 3, len(T[i]), len(T[i]) + 5, UMAX]) for i in range(n)], np.uint64)
    ps = A.PairSet.from_seqs(P, T)
    for sc in (ONT, A.Affine.HIFI):
        for clip in ((0, True), (25, True), (50, False), (IMAX, True), None):
            for use_geom, use_cut in ((False, False), (True, True), (False, True)):
                args = (ps, mode, sc, bands, geoms if use_geom else None, clip, cuts if use_cut else None)
                pass_policy(0)
                ref = A.affine_dp_batch(*args)
                pass_policy(2)
                got16 = A.affine_dp_batch(*args)
                pass_policy(1)
                got = A.affine_dp_batch(*args)
                for i in range(n):
                    assert tuple(got[i]) == tuple(ref[i]) == tuple(got16[i]), (i, len(P[i]), len(T[i]),
                                                                               int(bands[i]), sc, clip,
                                                                               use_geom, use_cut)
                for i in range(0, n, 17):
                    want = O.affine_dp(P[i], T[i], int(mode), (sc.mat, sc.mis, sc.open, sc.ext), int(bands[i]),
                                       int(geoms[i]) if use_geom else None, clip,
                                       int(cuts[i]) if use_cut else None)
                    w = (want.score, want.score_col, want.clip_best, want.clip_i, want.clip_j, want.text_read,
                         want.last_row, want.cut_best, want.cut_score)
                    assert tuple(got[i]) == w, (i, sc, clip, use_geom, use_cut)


def test_affine_dp_packed_pass_range_edges(pass_policy):
    # Reads around the packed pass's range limits for ONT at band 16 (three-term max below
    # 0x7C00: m near 5,200; 16 bits: m near 10,900; past it the 32-bit pass), a long match
    # then junk, so X-drop stops late or never; equal to the 32-bit pass and to the oracle.
    rng = random.Random(137)
    P, T = [], []
    for m in (5100, 5250, 5400, 10700, 10900, 11100):
        t = [rng.randrange(4) for _ in range(m + 40)]
        p = [x if rng.random() > 0.04 else rng.randrange(4) for x in t[:m - 200]] + \
            [rng.randrange(4) for _ in range(200)]
        P.append(p)
        T.append(t)
    ps = A.PairSet.from_seqs(P, T)
    for clip in ((50, True), (IMAX, True)):
        pass_policy(1)
        got = A.affine_dp_batch(ps, A.Mode.Prefix, ONT, 16, None, clip, None)
        pass_policy(0)
        ref = A.affine_dp_batch(ps, A.Mode.Prefix, ONT, 16, None, clip, None)
        for i in range(len(P)):
            assert tuple(got[i]) == tuple(ref[i]), (i, len(P[i]), clip)
            want = O.affine_dp(P[i], T[i], O.PREFIX, O.ONT, 16, None, clip, None)
            assert (got[i].score, got[i].clip_best, got[i].clip_i, got[i].clip_j, got[i].last_row) == \
                (want.score, want.clip_best, want.clip_i, want.clip_j, want.last_row), (i, clip)
