"""Pins the CPU oracle against the reference's own known answers and naive definitions.

Every expected value below is a literal from crates/map/src/align.rs's tests (cited per test) or
comes from a naive definition restated from those tests (full-matrix DPs in sf_oracle.c), which
are different formulas from the oracle's bit-parallel and prefix-max passes.
"""
import random

import pytest

from oracle import sf_oracle as O

INFIX, PREFIX, GLOBAL = O.INFIX, O.PREFIX, O.GLOBAL
ONT, HIFI = O.ONT, O.HIFI
IMAX = 2**31 - 1


# ---- align.rs:2908-2926 `examples` ----
def test_myers_examples():
    assert O.infix_distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3]) == 0
    assert O.infix_distance([0, 1, 2, 3], [0, 1, 3]) == 1
    assert O.infix_distance([0, 1, 2, 3], []) == 4
    assert O.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], PREFIX, 10) == 2
    assert O.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], GLOBAL, 10) == 3
    assert O.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], GLOBAL, 2) is None
    assert O.distance([], [0, 1], GLOBAL, 5) == 2


# ---- align.rs:3016-3043 `local_band_examples` ----
def test_local_band_examples():
    sc = ONT
    t = [3, 3, 0, 1, 2, 3, 0, 1, 2, 2, 2]
    p = [0, 1, 2, 3, 0, 1]
    assert O.local_band(p, t, sc, 2, 2, 5) == 12 + 10
    assert O.local_band(p, t, sc, -6, 11, 5) == 12 + 10
    assert O.local_band(p, t, sc, 0, 0, 0) == 0
    assert O.local_band(p, t, sc, 6, 6, 0) == 6
    junk = [2, 2, 2, 0, 1, 2, 3, 0, 1, 3, 3, 3]
    assert O.local_band(junk, t, sc, -12, 11, 5) == 12
    assert O.local_band([4, 4, 4], [0, 0, 0], sc, -3, 3, 0) == 0
    assert O.local_band([], t, sc, -3, 3, 5) == 0
    assert O.local_band(p, t, sc, 3, 2, 5) == 0
    assert O.local_band_ends(p, t, sc, 2, 2, (5, 0)) == 12 + 5
    assert O.local_band_ends(p, t, sc, 2, 2, (0, 3)) == 12 + 3
    assert O.local_band_ends([0, 1, 2, 2], [0, 1, 2], sc, 0, 0, (0, 5)) == 6
    assert O.local_band_ends([0, 1, 2, 2], [0, 1, 2], sc, 0, 0, (5, 0)) == 6 + 5


# ---- align.rs:3118-3129 `affine_extend_examples` ----
def test_affine_extend_examples():
    t = [0, 1, 2, 3, 0, 1, 2, 3]
    p = [0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3]
    assert O.affine_extend(p, t, ONT, None, 0, IMAX) == O.naive_extend(p, t, ONT)
    assert O.affine_extend([], t, ONT, None, 0, IMAX) == 0


# ---- align.rs:3131-3145 `affine_extend_end_examples` ----
def test_affine_extend_end_examples():
    def end(p, t, bonus):
        x, i, j, _ = O.affine_extend_window(p, t, len(t), ONT, 16, bonus, 50)
        return (x, i, j)

    t = [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
    assert end([0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3], t, 0) == (12, 6, 6)
    assert end([0, 1, 2, 0, 1, 2, 3], t, 5) == (14 - 6 + 5, 7, 8)
    assert end([3, 3, 3], [0, 0, 0], 0) == (0, 0, 0)


# ---- align.rs:3147-3166 `affine_extend_window_cuts` ----
def test_affine_extend_window_cuts():
    t = [((i * 5 + i // 4) % 4) for i in range(140)]
    p = t[:70] + [[3, 3, 0, 0][i % 4] for i in range(40)]
    r = O.affine_extend_window(p, t, len(t), ONT, 16, 5, 30)
    assert r[3] < len(t), r
    for cut in (r[3], r[3] + 1, len(t) - 1):
        assert O.affine_extend_window(p, t[:cut], len(t), ONT, 16, 5, 30)[0] == r[0]
    short = O.affine_extend_window(p, t[:40], len(t), ONT, 16, 5, 30)
    assert (short[2], short[3]) == (40, 40)


# ---- align.rs:3168-3183 `affine_extend_end_reads_a_prefix` ----
def test_affine_extend_end_reads_a_prefix():
    t = [((i * 7 + i // 3) % 4) for i in range(120)]
    p = t[:60] + [[3, 3, 0, 0][i % 4] for i in range(40)]
    r = O.affine_extend_window(p, t, len(t), ONT, 16, 5, 30)
    assert r[3] < len(t), r
    u = list(t)
    for q in range(r[3], len(u)):
        u[q] = 3 - u[q]
    assert O.affine_extend_window(p, u, len(u), ONT, 16, 5, 30) == r


# ---- align.rs:3185-3196 `affine_examples` ----
def test_affine_examples():
    p = [0, 1, 2, 3, 0, 1, 2, 3]
    assert O.affine(p, p, GLOBAL, ONT, None) == 16
    t = [0, 1, 2, 3, 0, 1, 2, 3, 2, 2]
    assert O.affine(p, t, GLOBAL, ONT, None) == 16 - 8
    assert O.affine(p, t, PREFIX, ONT, None) == 16
    assert O.affine([], [0, 1], GLOBAL, ONT, None) == -8
    assert O.affine([0, 1], [], GLOBAL, ONT, None) == -8


# ---- crates/map/proptest-regressions/align.txt: the shrunk inputs, replayed ----
def test_regression_random_vs_naive():
    # cc 5c23f3c9...: p = [0], t = 120 zeros, k = 0
    p, t = [0], [0] * 120
    for mode in (INFIX, PREFIX, GLOBAL):
        want = O.naive_edit(p, t, mode)
        assert O.distance(p, t, mode, 0) == (want if want <= 0 else None)


def test_regression_related_vs_naive():
    # cc 8b234219...: t as recorded, edits [(0,0,2),(0,0,0)], a = 0, len = 60, slack = 1
    t = ([0] * 36 + [1] + [0] * 72 + [2, 0, 1] + [0] * 84 + [1, 1, 1, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 2]
         + [0] * 72)
    p = list(t[0:60])
    for pos, b, kind in [(0, 0, 2), (0, 0, 0)]:
        i = pos % len(p)
        if kind == 0:
            p[i] = b
        elif kind == 1:
            p.insert(i, b)
        else:
            del p[i]
    d = O.naive_edit(p, t, INFIX)
    for k in (max(d - 1, 0), d, d + 1, O.NONE):
        for mode in (INFIX, PREFIX, GLOBAL):
            want = O.naive_edit(p, t, mode)
            assert O.distance(p, t, mode, k) == (want if want <= k else None)


def test_regression_local_band():
    # cc 97b72451...: p = [0], t = [], bonus = 1
    for sc in (ONT, HIFI):
        assert O.local_band_ends([0], [], sc, -1, 0, (1, 1)) == O.naive_local([0], [], sc, (1, 1))


def _rand_pair(n, div, seed):
    """align.rs:3503-3522 `rand_pair` (xorshift), restated for the recorded seeds."""
    x = seed
    M = (1 << 64) - 1

    def nxt():
        nonlocal x
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        return x

    t = [nxt() & 3 for _ in range(n)]
    p = []
    for b in t:
        r = nxt() % (3 * div)
        if r == 0:
            p.append((b + 1) & 3)
        elif r == 1:
            pass
        elif r == 2:
            p.extend([b, nxt() & 3])
        else:
            p.append(b)
    return p, t


@pytest.mark.parametrize("seed,n,div,band", [
    (16199839844444629902, 399, 8, 1),
    (2032331130425124312, 249, 21, 56),
    (3608202155028448350, 195, 12, 1),
    (4662852956731976402, 68, 10, 29),
])
def test_regression_narrow_inputs(seed, n, div, band):
    # The narrow_cells_match_wide regressions: the oracle is the 32-bit pass, whose banded scores
    # never exceed the exact ones and equal them once the band covers everything.
    p, t = _rand_pair(n, div, seed)
    for mode in (GLOBAL, PREFIX):
        exact = O.naive_affine(p, t, mode, ONT)
        assert O.affine(p, t, mode, ONT, band) <= exact
        assert O.affine(p, t, mode, ONT, 10_000) == exact


# ---- naive-definition properties (random_vs_naive, related_vs_naive, affine_vs_naive) ----
def _mutate(v, edits):
    v = list(v)
    for pos, b, kind in edits:
        if not v:
            break
        i = pos % len(v)
        if kind == 0:
            v[i] = b
        elif kind == 1:
            v.insert(i, b)
        else:
            del v[i]
    return v


def test_myers_vs_naive_random():
    rng = random.Random(1)
    for _ in range(250):
        p = [rng.randrange(4) for _ in range(rng.randrange(0, 140))]
        t = [rng.randrange(4) for _ in range(rng.randrange(0, 140))]
        k = rng.choice([O.NONE, rng.randrange(0, 200)])
        for mode in (INFIX, PREFIX, GLOBAL):
            want = O.naive_edit(p, t, mode)
            assert O.distance(p, t, mode, k) == (want if want <= k else None), (p, t, mode, k)


def test_myers_vs_naive_related():
    rng = random.Random(2)
    for _ in range(120):
        t = [rng.randrange(4) for _ in range(rng.randrange(100, 400))]
        a, ln = rng.randrange(50), rng.randrange(60, 350)
        end = min(a + ln, len(t))
        p = _mutate(t[min(a, end):end], [(rng.randrange(500), rng.randrange(4), rng.randrange(3))
                                         for _ in range(rng.randrange(30))])
        d = O.naive_edit(p, t, INFIX)
        for k in (max(d - 1, 0), d, d + rng.randrange(20), O.NONE):
            for mode in (INFIX, PREFIX, GLOBAL):
                want = O.naive_edit(p, t, mode)
                assert O.distance(p, t, mode, k) == (want if want <= k else None)


def test_batch_distances_match_single():
    # align.rs:3376-3432: the 8-lane union band gives each text's own distance_with result.
    rng = random.Random(3)
    for _ in range(60):
        base = [rng.randrange(4) for _ in range(rng.randrange(0, 500))]
        pa, plen = rng.randrange(300), rng.randrange(400)
        end = min(pa + plen, len(base))
        p = _mutate(base[min(pa, end):end], [(rng.randrange(700), rng.randrange(4), rng.randrange(3))
                                             for _ in range(rng.randrange(40))])
        rnd = [rng.randrange(4) for _ in range(rng.randrange(300))]
        ts = []
        for _ in range(rng.randrange(1, 20)):
            a, ln = rng.randrange(700), rng.randrange(700)
            if rng.random() < 0.3:
                ts.append(rnd[:min(ln, len(rnd))])
            else:
                e = min(a + ln, len(base))
                ts.append(_mutate(base[min(a, e):e], [(rng.randrange(700), rng.randrange(4), rng.randrange(3))
                                                      for _ in range(rng.randrange(25))]))
        for mode in (INFIX, PREFIX, GLOBAL):
            exact = [O.distance(p, t, mode, O.NONE) for t in ts]
            k = O.NONE if rng.random() < 0.3 else min(exact) + rng.randrange(60)
            want = [O.distance(p, t, mode, k) for t in ts]
            assert O.batch_distances(p, ts, mode, k) == want


def test_affine_vs_naive():
    # align.rs:3312-3341
    rng = random.Random(4)
    for it in range(150):
        t = [rng.randrange(4) for _ in range(rng.randrange(0, 160))]
        a, ln = rng.randrange(30), rng.randrange(150)
        end = min(a + ln, len(t))
        p = _mutate(t[min(a, end):end], [(rng.randrange(200), rng.randrange(4), rng.randrange(3))
                                         for _ in range(rng.randrange(20))])
        for mode in (GLOBAL, PREFIX, INFIX):
            want = O.naive_affine(p, t, mode, ONT)
            assert O.affine(p, t, mode, ONT, None) == want
            assert O.affine(p, t, mode, ONT, 8) <= want
            assert O.affine(p, t, mode, ONT, 400) == want
        if len(p) <= 25 and len(t) <= 30 and it % 5 == 0:
            assert O.affine_extend(p, t, ONT, None, 0, IMAX) == O.naive_extend(p, t, ONT)


def test_local_band_matches_definition():
    # align.rs:3048-3063
    rng = random.Random(5)
    for _ in range(60):
        p = [rng.randrange(4) for _ in range(rng.randrange(0, 7))]
        t = [rng.randrange(4) for _ in range(rng.randrange(0, 7))]
        bonus = (rng.randrange(8), rng.randrange(8))
        for sc in (ONT, HIFI):
            m, n = len(p), len(t)
            assert O.local_band_ends(p, t, sc, -m, n, bonus) == O.naive_local(p, t, sc, bonus)


def test_local_band_matches_matrix():
    # align.rs:3065-3115
    rng = random.Random(6)
    for _ in range(150):
        base = [rng.randrange(4) for _ in range(rng.randrange(1, 200))]
        p = base[rng.randrange(len(base)):]
        p = p[:max(len(p) - rng.randrange(60), 0)]
        p = _mutate(p, [(rng.randrange(300), rng.randrange(5), rng.randrange(3)) for _ in range(rng.randrange(25))])
        if rng.random() < 0.5:
            p = [rng.randrange(4) for _ in range(rng.randrange(20))] + p + [rng.randrange(4) for _ in range(rng.randrange(20))]
        dlo, width = rng.randrange(-250, 250), rng.randrange(150)
        bonus = (rng.randrange(8), rng.randrange(8))
        for sc in (ONT, HIFI):
            got = O.local_band_ends(p, base, sc, dlo, dlo + width, bonus)
            assert got == O.naive_local_band(p, base, sc, dlo, dlo + width, bonus)
            assert got <= O.local_band_ends(p, base, sc, dlo - 3, dlo + width + 5, bonus)


def test_window_cut_equals_cut_text():
    # align.rs:3295-3306: with no X-drop the cut score equals the pass over the cut text.
    rng = random.Random(7)
    for _ in range(60):
        base = [rng.randrange(4) for _ in range(rng.randrange(1, 150))]
        p = _mutate(base, [(rng.randrange(160), rng.randrange(4), rng.randrange(3)) for _ in range(rng.randrange(12))])
        ts = [base]
        for _ in range(rng.randrange(1, 5)):
            keep = rng.randrange(160)
            t = base[:keep] + [rng.randrange(4) for _ in range(rng.randrange(40))] + base[keep:]
            ts.append(t[:max(len(t) - rng.randrange(20), 0)])
        band, bonus = rng.randrange(2, 24), rng.randrange(8)
        w = max(len(t) for t in ts) + rng.randrange(40)
        c = min(len(t) for t in ts) + rng.randrange(20)
        for t in ts:
            both = O.affine_extend_cut_bonus(p, t, ONT, band, bonus, IMAX, False, w, c)
            want = O.affine_extend_window(p, t[:min(len(t), c)], w, ONT, band, bonus, IMAX)[0]
            assert both[1] == want
            whole = O.affine_extend_cut_bonus(p, t, ONT, band, bonus, IMAX, False, w, None)[0]
            assert both[0] == whole
            assert whole == O.affine_extend_window(p, t, w, ONT, band, bonus, IMAX)[0]


def test_affine_cells_counts_rows():
    # DP_WORK (align.rs:2573-2578) over a Global band-64 pass: rows 1..m of hi + 1 - lo.
    m, n, band = 300, 280, 64
    below, above = band + max(m - n, 0), band + max(n - m, 0)
    want = sum(min(i + above, n) + 1 - max(i - below, 0) for i in range(1, m + 1))
    assert O.affine_cells(m, n, GLOBAL, band) == want
    assert O.affine_cells(m, n, INFIX, band) == m * (n + 1)


@pytest.mark.skipif(not O.avx512_available(), reason="host without AVX-512 F/BW/VL")
def test_avx512_rows_equal_scalar_rows():
    # The CPU baseline's AVX-512 rows (run_core::<Avx512Rows>) against the scalar rows: rows
    # shorter than a vector, whole vectors and ragged tails, every mode, bands from 0 to exact.
    import numpy as np
    rng = random.Random(29)
    P, T, B = [], [], []
    for _ in range(600):
        n = rng.choice([0, 1, 5, 15, 16, 17, 31, 33, 64, rng.randrange(0, 400)])
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.12 else rng.randrange(4) for x in t]
        for _ in range(rng.randrange(0, 4)):
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(0, 6))] + p[k + rng.randrange(0, 6):]
        P.append(bytes(p))
        T.append(bytes(t))
        B.append(rng.choice([0, 1, 3, 15, 16, 17, 40, 100, O.UMAX]))
    poff = np.cumsum([0] + [len(p) for p in P[:-1]]).astype(np.uint64)
    toff = np.cumsum([0] + [len(t) for t in T[:-1]]).astype(np.uint64)
    pat = np.frombuffer(b"".join(P) + b"\0", np.uint8)
    txt = np.frombuffer(b"".join(T) + b"\0", np.uint8)
    plen = np.array([len(p) for p in P], np.uint32)
    tlen = np.array([len(t) for t in T], np.uint32)
    band = np.array(B, np.uint64)
    for mode in (INFIX, PREFIX, GLOBAL):
        for sc in (ONT, HIFI, (1, 1, 0, 1), (3, 2, 5, 1)):
            a = O.batch_affine(pat, poff, plen, txt, toff, tlen, mode, sc, band, threads=4)
            b = O.batch_affine(pat, poff, plen, txt, toff, tlen, mode, sc, band, threads=4, avx512=True)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (mode, sc)
