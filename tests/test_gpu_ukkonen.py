"""GPU parity of the Myers passes over Ukkonen's band (sf_set_myers_ukk): distances, end and start
columns bit for bit against the oracle (the reference's `distance_with`, Ukkonen cut-off included,
align.rs:92-168), and equal to the passes that compute every block, at thresholds from 0 to past
the patterns' lengths, in all three modes.

Each wave mixes patterns of different block counts (65-1,536 bases: every lane class that runs the
band), texts that hold the pattern with few errors, with many, unrelated texts (distance far above
k, so every lane's band collapses) and short or empty texts.
"""
import random

import pytest

from oracle import sf_oracle as O
from spearfish_amd import _native as N
from spearfish_amd import align as A

pytestmark = pytest.mark.gpu

MODES = (A.Mode.Infix, A.Mode.Prefix, A.Mode.Global)
U32 = A.U32_MAX


def opt(x):
    return U32 if x is None else x


def mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 2:
            out.append(rng.randrange(4))  # substitution
        elif r < 3 * rate / 4:
            out.extend([c, rng.randrange(4)])  # insertion
        elif r < rate:
            pass  # deletion
        else:
            out.append(c)
    return out


def pairs(seed, n):
    rng = random.Random(seed)
    P, T = [], []
    for i in range(n):
        m = rng.choice([65, 100, 128, 129, 192, 193, 300, 512, 700, 960, 1000, 1024, 1025, 1100, 1280, 1281,
                        1400, 1535, 1536, rng.randrange(65, 1537)])
        p = [rng.randrange(4) for _ in range(m)]
        kind = i % 5
        if kind == 0:    # the pattern with 6% errors inside a longer text
            t = [rng.randrange(4) for _ in range(rng.randrange(0, 200))] + mutate(rng, p, 0.06) + \
                [rng.randrange(4) for _ in range(rng.randrange(0, 200))]
        elif kind == 1:  # 20% errors, whole
            t = mutate(rng, p, 0.20)
        elif kind == 2:  # unrelated
            t = [rng.randrange(4) for _ in range(rng.randrange(1, 2 * m))]
        elif kind == 3:  # short or empty
            t = [rng.randrange(4) for _ in range(rng.randrange(0, 40))]
        else:            # a prefix of the pattern with errors, then more text
            t = mutate(rng, p[:rng.randrange(1, m + 1)], 0.03) + [rng.randrange(4) for _ in range(rng.randrange(0, 300))]
        P.append(p)
        T.append(t)
    return A.PairSet.from_seqs(P, T)


@pytest.fixture
def ukk():
    lib = N.load()
    N.check(lib.sf_set_myers_ukk(1))
    yield lib
    N.check(lib.sf_set_myers_ukk(1))


@pytest.mark.parametrize("mode", MODES)
def test_band_equals_oracle_and_full_columns(ukk, mode):
    ps = pairs(7 + int(mode), 640)
    for k in (0, 1, 7, 30, 63, 64, 65, 100, 200, 500, 1000, 1600, U32):
        N.check(ukk.sf_set_myers_ukk(1))
        dist, end, start = A.myers_batch(ps, mode, k, locate=2)
        N.check(ukk.sf_set_myers_ukk(0))
        full = A.myers_batch(ps, mode, k, locate=2)
        assert (dist == full[0]).all() and (end == full[1]).all() and (start == full[2]).all(), (mode, k)
        for i in range(0, len(ps), 3 if k != U32 else 9):
            p, t = ps.pattern(i), ps.text(i)
            d, e, s = O.locate(p, t, int(mode), k)
            assert (int(dist[i]), int(end[i]), int(start[i])) == (opt(d), opt(e), opt(s)), (i, len(p), len(t), mode, k)


def test_distance_exactly_at_the_threshold(ukk):
    # k = the true distance (exact, kept) and k = distance - 1 (None): the band's edge cases.
    ps = pairs(99, 256)
    dist, _, _ = A.myers_batch(ps, A.Mode.Infix, U32, locate=0)
    for i in range(0, len(ps), 8):
        p, t = ps.pattern(i), ps.text(i)
        d = int(dist[i])
        one = A.PairSet.from_seqs([p], [t])
        assert int(A.myers_batch(one, A.Mode.Infix, d, locate=0)[0][0]) == d
        if d > 0:
            assert int(A.myers_batch(one, A.Mode.Infix, d - 1, locate=0)[0][0]) == U32
    # The whole set at per-pair thresholds is not an API; one threshold per call across the set:
    for k in sorted({int(x) for x in dist[:64] if x != U32})[:6]:
        got = A.myers_batch(ps, A.Mode.Infix, k, locate=1)
        for i in range(len(ps)):
            want = O.locate(ps.pattern(i), ps.text(i), int(A.Mode.Infix), k)
            assert (int(got[0][i]), int(got[1][i])) == (opt(want[0]), opt(want[1])), (i, k)


def long_pairs(seed, n):
    rng = random.Random(seed)
    P, T = [], []
    for i in range(n):
        m = rng.choice([1537, 1600, 2047, 2048, 2049, 3071, 3072, 3073, 4100, 6000, 9000, 12000,
                        rng.randrange(1537, 12001)])
        p = [rng.randrange(4) for _ in range(m)]
        kind = i % 6
        flank = lambda: [rng.randrange(4) for _ in range(rng.randrange(0, 400))]
        if kind == 0:    # exact, inside a longer text (d = 0)
            t = flank() + list(p) + flank()
        elif kind == 1:  # 5% errors
            t = flank() + mutate(rng, p, 0.05) + flank()
        elif kind == 2:  # 12%, as the long workload's reads
            t = mutate(rng, p, 0.12) + flank()
        elif kind == 3:  # 40%: windows wider than a stripe
            t = flank() + mutate(rng, p, 0.40)
        elif kind == 4:  # unrelated (d far above 1,024: stripes start at column 1)
            t = [rng.randrange(4) for _ in range(rng.randrange(m // 2, 2 * m))]
        else:            # a piece of the pattern only
            lo = rng.randrange(0, m // 2)
            t = mutate(rng, p[lo:lo + rng.randrange(1, m // 2)], 0.03)
        P.append(p)
        T.append(t)
    return A.PairSet.from_seqs(P, T)


def test_start_over_the_diagonal_band_long_patterns(ukk):
    # Patterns over 1,536 bases find `start` stripe by stripe over the band |i - j| <= d
    # (myers_start.hip); against the oracle and the pipelines' full reversed pass.
    ps = long_pairs(123, 240)
    N.check(ukk.sf_set_myers_ukk(1))
    dist, end, start = A.myers_batch(ps, A.Mode.Infix, U32, locate=2)
    N.check(ukk.sf_set_myers_ukk(0))
    full = A.myers_batch(ps, A.Mode.Infix, U32, locate=2)
    assert (dist == full[0]).all() and (end == full[1]).all()
    assert (start == full[2]).all(), [i for i in range(len(ps)) if start[i] != full[2][i]][:10]
    for i in range(0, len(ps), 2):
        p, t = ps.pattern(i), ps.text(i)
        d, e, s = O.locate(p, t, int(A.Mode.Infix))
        assert (int(dist[i]), int(end[i]), int(start[i])) == (opt(d), opt(e), opt(s)), (i, len(p), len(t))
