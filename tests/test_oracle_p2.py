"""The oracle's passes with per-row bookkeeping against an independent full-matrix restatement
(tests/fullmatrix.py): every `Dp` field (align.rs:2111-2129) of `affine_dp` with finite X-drops
(the stop row, `text_read`, the first maximum of the row that raised the best as the clip cell),
cuts (`cut_best`, `cut_score`) and band layouts (`geom`), and the wrappers `affine_extend_window`
and `affine_extend_sorted_cut_bonus` with `diagonal`, `window`, `track_cut` and
`bonus_within_cut`, on random small inputs.

Then each of several plausible one-token mistakes is planted in a copy of the oracle's C source,
compiled here, and shown to fail these comparisons: the checks pin those branches.
"""
import ctypes as C
import os
import random
import subprocess

import pytest

import fullmatrix as FM
from oracle import sf_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCORINGS = [O.ONT, O.HIFI, (1, 1, 0, 1), (3, 2, 5, 1), (2, 3, 1, 1)]
XDROPS = [0, 1, 2, 3, 5, 8, 13, 2**31 - 1]


def related(rng, n):
    """A random sequence and a mutated copy of part of it, so alignments run long and then fail."""
    base = [rng.randrange(4) for _ in range(n)]
    v = list(base[rng.randrange(0, max(1, n // 3)):])
    for _ in range(rng.randrange(0, 8)):
        if not v:
            break
        i = rng.randrange(len(v))
        k = rng.randrange(3)
        if k == 0:
            v[i] = rng.randrange(4)
        elif k == 1:
            v.insert(i, rng.randrange(4))
        else:
            del v[i]
    if rng.random() < 0.3:  # junk tail: X-drop territory
        v = v[: len(v) // 2] + [rng.randrange(4) for _ in range(rng.randrange(0, 20))]
    return base, v


def dp_cases(seed, count):
    rng = random.Random(seed)
    for _ in range(count):
        a, b = related(rng, rng.randrange(0, 28))
        p, t = (a, b) if rng.random() < 0.5 else (b, a)
        mode = rng.choice([O.PREFIX, O.PREFIX, O.GLOBAL, O.INFIX])
        band = rng.choice([0, 1, 2, 3, 5, 8, None])
        geom = len(t) if (mode == O.GLOBAL or rng.random() < 0.4) else rng.randrange(0, 2 * len(t) + 6)
        clip = (rng.choice(XDROPS), rng.random() < 0.7) if rng.random() < 0.8 else None
        cut = rng.choice([None, rng.randrange(0, len(t) + 3)])
        yield p, t, mode, rng.choice(SCORINGS), band, geom, clip, cut


def check_dp(seed=11, count=1500):
    bad = 0
    for p, t, mode, sc, band, geom, clip, cut in dp_cases(seed, count):
        want = FM.dp(p, t, mode, sc, band, geom, clip, cut)
        d = O.affine_dp(p, t, mode, sc, band, geom, clip, cut)
        got = FM.Dp(d.score, d.score_col, d.clip_best, d.clip_i, d.clip_j, d.text_read, d.last_row, d.cut_best,
                    d.cut_score)
        if got != want:
            bad += 1
    return bad


def wrapper_cases(seed, count):
    rng = random.Random(seed)
    for _ in range(count):
        p, t = related(rng, rng.randrange(1, 30))
        if rng.random() < 0.5:
            t = t + [rng.randrange(4) for _ in range(rng.randrange(0, 12))]
        sc = rng.choice(SCORINGS)
        band = rng.choice([0, 1, 2, 4, 8])
        bonus = rng.choice([0, 1, 5])
        xdrop = rng.choice(XDROPS)
        window = rng.randrange(0, len(t) + 10)
        cut = rng.choice([None, rng.randrange(0, len(t) + 2)])
        yield p, t, sc, band, bonus, xdrop, window, cut, rng.random() < 0.5, rng.random() < 0.7, rng.random() < 0.5


def check_wrappers(seed=12, count=1500):
    bad = 0
    for p, t, sc, band, bonus, xdrop, window, cut, diagonal, track, within in wrapper_cases(seed, count):
        want = FM.extend_cut_bonus(p, t, sc, band, bonus, xdrop, diagonal, window, cut, track, within)
        got = O.affine_extend_cut_bonus(p, t, sc, band, bonus, xdrop, diagonal, window, cut, track, within)
        bad += tuple(got) != want
        want = FM.extend_window(p, t, max(window, len(t)), sc, band, bonus, xdrop)
        got = O.affine_extend_window(p, t, max(window, len(t)), sc, band, bonus, xdrop)
        bad += tuple(got) != want
        bad += O.affine_extend(p, t, sc, band, bonus, xdrop) != FM.extend(p, t, sc, band, bonus, xdrop)
    return bad


def test_affine_dp_every_field_matches_the_full_matrix():
    assert check_dp() == 0


def test_extension_wrappers_match_the_full_matrix():
    assert check_wrappers() == 0


def test_full_matrix_restatement_against_the_reference_examples():
    # align.rs:3131-3145 `affine_extend_end_examples` (band 16, X-drop 50; the window form with the
    # text's own length), and align.rs:3118-3129 `affine_extend_examples`, through the independent
    # restatement itself.
    s = O.ONT
    t = [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
    end = lambda p, t, bonus: FM.extend_window(p, t, len(t), s, 16, bonus, 50)[:3]
    assert end([0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3], t, 0) == (12, 6, 6)
    assert end([0, 1, 2, 0, 1, 2, 3], t, 5) == (14 - 6 + 5, 7, 8)
    assert end([3, 3, 3], [0, 0, 0], 0) == (0, 0, 0)
    assert FM.extend([], [0, 1, 2, 3, 0, 1, 2, 3], s, None, 0, 2**31 - 1) == 0
    p, t8 = [0, 1, 2, 3, 0, 1, 3, 3, 3, 3, 3], [0, 1, 2, 3, 0, 1, 2, 3]
    assert FM.extend(p, t8, s, None, 0, 2**31 - 1) == O.naive_extend(p, t8, s)


# One-token mistakes planted in copies of oracle/sf_oracle.c: (what, original, mutant).
MUTANTS = [
    ("X-drop stop test < as <=", "if (row < sat_sub_i(clip_best, xdrop)) {", "if (row <= sat_sub_i(clip_best, xdrop)) {"),
    ("diagonal band layout min as max", "uint64_t geom = diagonal ? (m < n ? m : n)", "uint64_t geom = diagonal ? (m > n ? m : n)"),
    ("bonus_within_cut ignored", "if (bonus_within_cut && cut != SFO_UMAX) end =", "if (0 && bonus_within_cut && cut != SFO_UMAX) end ="),
    ("text_read at a stop as the row's lo", "sfo_dp r = {NEG, 0, clip_best, ci, cj, hi, i, cut_best, NEG};",
     "sfo_dp r = {NEG, 0, clip_best, ci, cj, lo, i, cut_best, NEG};"),
    ("best cell on a tie moves to the later row", "if (row > clip_best) {\n                clip_best = row;",
     "if (row >= clip_best) {\n                clip_best = row;"),
    ("cut row stops one column short", "for (uint64_t j = lo; j <= cut; j++) row = max32(row, nh[j]);",
     "for (uint64_t j = lo; j < cut; j++) row = max32(row, nh[j]);"),
    ("cut score over the whole last row", "uint64_t top = phi < cut ? phi : cut;", "uint64_t top = phi;"),
]


@pytest.mark.parametrize("what,orig,mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_a_planted_mistake_fails_the_checks(tmp_path, what, orig, mutant):
    src = open(os.path.join(ROOT, "oracle", "sf_oracle.c")).read()
    assert src.count(orig) == 1, what
    path = tmp_path / "mutant.c"
    path.write_text(src.replace(orig, mutant))
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-fwrapv", "-std=c11", "-shared", "-I", os.path.join(ROOT, "oracle"),
                    "-o", str(so), str(path), "-lpthread"], check=True)
    saved = O._lib
    O._lib = O.declare(C.CDLL(str(so)))
    try:
        bad = check_dp(count=600) + check_wrappers(count=600)
    finally:
        O._lib = saved
    assert bad > 0, what
