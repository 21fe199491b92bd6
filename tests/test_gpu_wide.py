"""Bands wider than 2,048 cells a row (the wide-band kernel, affine_wide.hip) against the oracle:
score runs in all three modes (Infix is unbanded: m + n + 1 cells a row), passes with X-drop, best
cell, cut and band layout (align.rs:2446-2666), and local passes over thousands of diagonals
(align.rs:783-852), mixed in one batch with narrow pairs that take the register-resident kernels."""
import random

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A

pytestmark = pytest.mark.gpu


def mutated(rng, n, rate=0.08):
    t = [rng.randrange(4) for _ in range(n)]
    p = []
    for c in t:
        r = rng.random()
        if r < rate / 3:
            continue
        if r < 2 * rate / 3:
            p.append(rng.randrange(4))
        p.append((c + 1) % 4 if r < rate else c)
    return p, t


def pairs(seed, sizes):
    rng = random.Random(seed)
    P, T = [], []
    for m in sizes:
        p, t = mutated(rng, m)
        if rng.random() < 0.5:
            t = [rng.randrange(4) for _ in range(rng.randrange(0, 300))] + t  # the read inside a longer text
        P.append(p)
        T.append(t)
    return P, T


SIZES = [3000, 2600, 2100, 1200, 900, 60, 5, 0, 2500]


@pytest.mark.parametrize("mode", [A.Mode.Infix, A.Mode.Prefix, A.Mode.Global])
def test_wide_scores_all_modes(mode):
    P, T = pairs(1 + int(mode), SIZES)
    ps = A.PairSet.from_seqs(P, T)
    # Bands: exact (m + n + 1 wide), 1,100 (over 2,200 wide), and narrow ones.
    bands = np.array([A.USIZE_MAX, 1100, 1500, 64, A.USIZE_MAX, 16, 3, 5, 1030], np.uint64)
    r = A.affine_batch(ps, mode, A.Affine.ONT, bands)
    for i in range(len(ps)):
        b = None if bands[i] == A.USIZE_MAX else int(bands[i])
        dp = O.affine_dp(P[i], T[i], int(mode), O.ONT, b)
        assert (int(r["score"][i]), int(r["score_col"][i])) == (dp.score, dp.score_col), (i, len(P[i]), len(T[i]), b)


def test_wide_passes_every_dp_field():
    rng = random.Random(9)
    P, T = pairs(5, [2200, 2500, 1800, 400, 2600])
    ps = A.PairSet.from_seqs(P, T)
    for mode in (A.Mode.Prefix, A.Mode.Global):
        for clip in (None, (50, True), (5, True), (2**31 - 1, False)):
            band = 1100
            geom = [len(t) if mode == A.Mode.Global else max(len(t), rng.randrange(0, 3000)) for t in T]
            cut = [rng.randrange(0, len(t) + 2) for t in T]
            got = A.affine_dp_batch(ps, mode, A.Affine.ONT, band, geom, clip, cut)
            for i in range(len(ps)):
                want = O.affine_dp(P[i], T[i], int(mode), O.ONT, band, geom[i], clip, cut[i])
                w = (want.score, want.score_col, want.clip_best, want.clip_i, want.clip_j, want.text_read,
                     want.last_row, want.cut_best, want.cut_score)
                assert tuple(got[i]) == w, (mode, clip, i)


def test_wide_local_bands():
    P, T = pairs(7, [2500, 3000, 900, 1800])
    ps = A.PairSet.from_seqs(P, T)
    for dlo, dhi in ((-1500, 1500), (-100, 2600), (-3000, -10), (20, 2200)):
        got = A.local_band_ends_batch(ps, A.Affine.ONT, dlo, dhi, (5, 7))
        for i in range(len(ps)):
            assert int(got[i]) == O.local_band_ends(P[i], T[i], O.ONT, dlo, dhi, (5, 7)), (dlo, dhi, i)
