"""A full-matrix restatement of the reference's banded affine pass with its per-row bookkeeping
(crates/map/src/align.rs:2446-2666 `run_core`: X-drop stop, best cell, cut tracking, band layout by
`geom`) and of the wrappers built on it (align.rs:710-729, 903-949, 957-982), written for the tests
from the definitions, independently of the oracle's C code.

It differs from the oracle in how it computes, not in what: every cell of the band gets textbook
Gotoh H, E and F (E from the cell on the left, F from the cell above, H the best of the diagonal,
E and F), out-of-band cells are minus infinity, and each row's bookkeeping is read off the finished
row.  The oracle (and the reference) compute a row in three passes, with the horizontal gaps as a
running prefix maximum, and track the row maximum as the row is written.  Small inputs only.

Global mode is restated only where the last row's band reaches the text's end (always so when the
band is laid out for the text itself); the reference reads its row buffer at column n otherwise.
"""
from __future__ import annotations

from typing import NamedTuple, Optional, Tuple

NEG = -(2**31) // 4
I32_MIN = -(2**31)
INFIX, PREFIX, GLOBAL = 0, 1, 2
_NINF = float("-inf")


class Dp(NamedTuple):
    score: int
    score_col: int
    clip_best: int
    clip_i: int
    clip_j: int
    text_read: int
    last_row: int
    cut_best: int
    cut_score: int


def sat_sub(a: int, b: int) -> int:
    return max(a - b, I32_MIN)


def sat_add(a: int, b: int) -> int:
    return min(a + b, 2**31 - 1)


def dp(p, t, mode: int, sc: Tuple[int, int, int, int], band: Optional[int] = None, geom: Optional[int] = None,
       clip: Optional[Tuple[int, bool]] = None, cut: Optional[int] = None) -> Dp:
    mat, mis, open_, ext = sc
    go = open_ + ext
    m, n = len(p), len(t)
    g = n if geom is None else geom
    full = mode == INFIX or band is None
    below = 0 if full else band + max(m - g, 0)
    above = 0 if full else band + max(g - m, 0)

    def rng(i):
        return (0, n) if full else (max(i - below, 0), min(i + above, n))

    H, E, F = {}, {}, {}
    lo, hi = rng(0)
    for j in range(lo, hi + 1):
        H[0, j] = 0 if (j == 0 or mode == INFIX) else -(open_ + ext * j)
    clip_best, clip_cell, cut_best = 0, (0, 0), 0
    last = (lo, hi)
    for i in range(1, m + 1):
        lo, hi = rng(i)
        if lo > hi:
            return Dp(NEG, 0, clip_best, clip_cell[0], clip_cell[1], n, i, cut_best, NEG)
        for j in range(lo, hi + 1):
            f = max(H.get((i - 1, j), _NINF) - go, F.get((i - 1, j), _NINF) - ext)
            if j == 0:  # column 0: only a vertical gap reaches it
                e, h = _NINF, f
            else:
                e = max(H.get((i, j - 1), _NINF) - go, E.get((i, j - 1), _NINF) - ext)
                s = mat if (p[i - 1] & 3) == (t[j - 1] & 3) else -mis
                h = max(H.get((i - 1, j - 1), _NINF) + s, e, f)
            H[i, j], E[i, j], F[i, j] = h, e, f
        row = [H[i, j] for j in range(lo, hi + 1)]
        row_max = max(row)
        if cut is not None and lo <= cut:
            cut_best = max(cut_best, max(row[: min(hi, cut) - lo + 1]))
        if clip is not None:
            xdrop, cell = clip
            if row_max > clip_best:
                clip_best = row_max
                if cell:
                    clip_cell = (i, lo + row.index(row_max))  # the row's first maximum
            if row_max < sat_sub(clip_best, xdrop):
                return Dp(NEG, 0, clip_best, clip_cell[0], clip_cell[1], hi, i, cut_best, NEG)
        last = (lo, hi)
    lo, hi = last
    row = [H[m, j] for j in range(lo, hi + 1)]
    if mode == GLOBAL:
        assert hi == n, "Global with a band that ends before the text is not restated"
        score, col = H[m, n], n
    else:
        score = max(row)
        col = lo + row.index(score)
    cut_score = max(row[: min(hi, cut) - lo + 1]) if (cut is not None and lo <= cut) else NEG
    return Dp(int(score), col, int(clip_best), clip_cell[0], clip_cell[1], hi, m, int(cut_best), int(cut_score))


def extend(p, t, sc, band, end_bonus, xdrop) -> int:
    """align.rs:710-729 `affine_extend`."""
    d = dp(p, t, PREFIX, sc, band, len(t), (xdrop, False))
    return max(d.clip_best, sat_add(d.score, end_bonus))


def extend_window(p, t, window, sc, band, end_bonus, xdrop):
    """align.rs:957-982 `affine_extend_window`: (score, pattern bases, text bases, text bases read)."""
    d = dp(p, t, PREFIX, sc, band, max(window, len(t)), (xdrop, True))
    whole = sat_add(d.score, end_bonus)
    if d.clip_best >= whole:
        return (d.clip_best, d.clip_i, d.clip_j, d.text_read)
    return (whole, len(p), d.score_col, d.text_read)


def extend_cut_bonus(p, t, sc, band, end_bonus, xdrop, diagonal, window, cut, track_cut, bonus_within_cut):
    """align.rs:903-949 `affine_extend_sorted_cut_bonus` for one text: (whole, cut, text bases)."""
    m, n = len(p), len(t)
    geom = min(m, n) if diagonal else max(n, window)
    d = dp(p, t, PREFIX, sc, band, geom, (xdrop, track_cut and cut is not None), cut)
    if bonus_within_cut and cut is not None:
        end = max(d.score, sat_add(d.cut_score, end_bonus))
    else:
        end = sat_add(d.score, end_bonus)
    whole = max(d.clip_best, end)
    cols = d.clip_j if d.clip_best >= end else d.score_col
    cutv = whole if cut is None else max(d.cut_best, sat_add(d.cut_score, end_bonus))
    return (whole, cutv, cols)
