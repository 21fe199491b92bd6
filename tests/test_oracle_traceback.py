"""Pins the oracle's own (reference-less) definitions: the affine CIGAR and the Myers end/start.

The reference has no traceback and no Myers end position (crates/map/src/align.rs returns scores
and distances only), so these rules are this project's.  They are pinned here by invariants
(rescoring, base consumption, band containment) and by an independent full-matrix Gotoh walk with
the same tie rules written over the textbook recurrences, which checks the prefix-max
formulation of E the oracle shares with the reference (align.rs:1060-1063).
"""
import random

from oracle import sf_oracle as O

INFIX, PREFIX, GLOBAL = O.INFIX, O.PREFIX, O.GLOBAL
NEG = O.NEG


def rescore(p, t, tstart, cig, sc):
    mat, mis, opn, ext = sc
    i, j, s = 0, tstart, 0
    for c in cig:
        ln, op = c >> 4, c & 15
        assert ln > 0
        if op == 7:
            for q in range(ln):
                assert p[i + q] == t[j + q]
            s += mat * ln
            i += ln
            j += ln
        elif op == 8:
            for q in range(ln):
                assert p[i + q] != t[j + q]
            s -= mis * ln
            i += ln
            j += ln
        elif op == 1:
            s -= opn + ext * ln
            i += ln
        elif op == 2:
            s -= opn + ext * ln
            j += ln
        else:
            raise AssertionError(op)
    return s, i, j


def band_range(m, n, mode, band, i):
    if mode == INFIX or band is None:
        return 0, n
    below, above = band + max(m - n, 0), band + max(n - m, 0)
    return max(i - below, 0), min(i + above, n)


def path_cells(tstart, cig, end_col):
    i, j = 0, tstart
    cells = [(0, j)]
    for c in cig:
        ln, op = c >> 4, c & 15
        for _ in range(ln):
            if op in (7, 8):
                i, j = i + 1, j + 1
            elif op == 1:
                i += 1
            else:
                j += 1
            cells.append((i, j))
    assert j == end_col
    return cells


def gotoh_walk(p, t, mode, sc, band):
    """Textbook banded Gotoh (E[i][j] = max(H[i][j-1] - go, E[i][j-1] - ext)), then the tie
    rules of the oracle's walk, written independently over full matrices."""
    mat, mis, opn, ext = sc
    go = opn + ext
    m, n = len(p), len(t)
    H = [[NEG] * (n + 1) for _ in range(m + 1)]
    E = [[NEG] * (n + 1) for _ in range(m + 1)]
    F = [[NEG] * (n + 1) for _ in range(m + 1)]
    inb = lambda i, j: band_range(m, n, mode, band, i)[0] <= j <= band_range(m, n, mode, band, i)[1]
    for j in range(n + 1):
        if inb(0, j):
            H[0][j] = 0 if (j == 0 or mode == INFIX) else -(opn + ext * j)
    for i in range(1, m + 1):
        lo, hi = band_range(m, n, mode, band, i)
        if lo > hi:
            return None
        for j in range(lo, hi + 1):
            if j == 0:
                H[i][0] = F[i][0] = max(H[i - 1][0] - go, F[i - 1][0] - ext)
                continue
            F[i][j] = max(H[i - 1][j] - go, F[i - 1][j] - ext, NEG)
            E[i][j] = max(H[i][j - 1] - go, E[i][j - 1] - ext)
            s = mat if p[i - 1] == t[j - 1] else -mis
            H[i][j] = max(H[i - 1][j - 1] + s, F[i][j], E[i][j])
    lo, hi = band_range(m, n, mode, band, m)
    if mode == GLOBAL:
        col = n
    else:
        col = max(range(lo, hi + 1), key=lambda j: (H[m][j], -j))
    score = H[m][col]
    ops = []
    i, j, st = m, col, "H"
    tstart = 0
    while True:
        if st == "H":
            if i == 0:
                if mode == INFIX:
                    tstart = j
                else:
                    ops += ["D"] * j
                break
            if j == 0:
                ops += ["I"] * i
                break
            s = mat if p[i - 1] == t[j - 1] else -mis
            if H[i][j] == H[i - 1][j - 1] + s:
                ops.append("=" if s == mat else "X")
                i, j = i - 1, j - 1
            elif H[i][j] == F[i][j]:
                st = "F"
            else:
                assert H[i][j] == E[i][j]
                st = "E"
        elif st == "F":
            ops.append("I")
            st = "F" if F[i][j] > H[i - 1][j] - go else "H"
            i -= 1
        else:
            ops.append("D")
            st = "E" if E[i][j] > H[i][j - 1] - go else "H"
            j -= 1
    ops.reverse()
    code = {"=": 7, "X": 8, "I": 1, "D": 2}
    cig = []
    for o in ops:
        if cig and (cig[-1] & 15) == code[o]:
            cig[-1] += 16
        else:
            cig.append(16 | code[o])
    return score, col, tstart, cig


def _related(rng, n_lo, n_hi, err):
    t = [rng.randrange(4) for _ in range(rng.randrange(n_lo, n_hi))]
    p = []
    for b in t:
        u = rng.random()
        if u < err / 2:
            p.append((b + 1 + rng.randrange(3)) & 3)
        elif u < 3 * err / 4:
            p.extend([b, rng.randrange(4)])
        elif u < err:
            pass
        else:
            p.append(b)
    return p, t


def test_cigar_invariants_and_independent_walk():
    rng = random.Random(11)
    for it in range(300):
        p, t = _related(rng, 0, 70, rng.choice([0.05, 0.15, 0.4]))
        if rng.random() < 0.15:
            p = [rng.randrange(4) for _ in range(rng.randrange(0, 40))]
        sc = rng.choice([O.ONT, O.HIFI, (1, 1, 0, 1), (3, 2, 5, 1)])
        mode = rng.choice([GLOBAL, PREFIX, INFIX])
        band = rng.choice([None, 0, 1, 3, 8, 30])
        r = O.align(p, t, mode, sc, band)
        dp = O.affine_dp(p, t, mode, sc, band)
        if r is None:
            assert dp.score == NEG
            continue
        score, col, tstart, cig = r
        assert score == dp.score and col == dp.score_col
        s, i, j = rescore(p, t, tstart, cig, sc)
        assert (s, i, j) == (score, len(p), col), (p, t, mode, sc, band, O.cigar_str(cig))
        if mode != INFIX:
            assert tstart == 0
        for (ci, cj) in path_cells(tstart, cig, col):
            lo, hi = band_range(len(p), len(t), mode, band, ci)
            assert lo <= cj <= hi
        want = gotoh_walk(p, t, mode, sc, band)
        assert want == (score, col, tstart, cig), (p, t, mode, sc, band)


def test_cigar_matches_naive_optimum_unbanded():
    rng = random.Random(12)
    for _ in range(200):
        p, t = _related(rng, 1, 60, 0.2)
        for mode in (GLOBAL, PREFIX, INFIX):
            score, col, tstart, cig = O.align(p, t, mode, O.ONT, None)
            assert score == O.naive_affine(p, t, mode, O.ONT)
            assert rescore(p, t, tstart, cig, O.ONT)[0] == score


def test_cigar_edge_cases():
    assert O.align([], [], GLOBAL) == (0, 0, 0, [])
    assert O.align([], [0, 1, 2], GLOBAL) == (-(4 + 2 * 3), 3, 0, [(3 << 4) | 2])
    assert O.align([0, 1], [], GLOBAL) == (-(4 + 2 * 2), 0, 0, [(2 << 4) | 1])
    assert O.align([0, 1, 2, 3], [0, 1, 2, 3], GLOBAL, band=0) == (8, 4, 0, [(4 << 4) | 7])
    assert O.align([], [0, 1], PREFIX) == (0, 0, 0, [])


def test_locate_brute_force():
    rng = random.Random(13)
    for _ in range(300):
        p, t = _related(rng, 0, 80, rng.choice([0.05, 0.2]))
        if rng.random() < 0.5 and len(t) > 10:
            t = [rng.randrange(4) for _ in range(rng.randrange(20))] + t + [rng.randrange(4) for _ in range(20)]
        mode = rng.choice([INFIX, PREFIX, GLOBAL])
        k = rng.choice([O.NONE, rng.randrange(0, 30)])
        d, e, s = O.locate(p, t, mode, k)
        want = O.naive_edit(p, t, mode)
        if want > k:
            assert (d, e, s) == (None, None, None)
            continue
        assert d == want
        row = O.naive_edit_lastrow(p, t, mode)
        if mode == GLOBAL:
            assert (e, s) == (len(t), 0)
            continue
        assert e == min(j for j in range(len(t) + 1) if row[j] == d)
        if mode == PREFIX:
            assert s == 0
        else:
            rrow = O.naive_edit_lastrow(p[::-1], t[:e][::-1], PREFIX)
            L = min(j for j in range(e + 1) if rrow[j] == d)
            assert s == e - L
        assert O.naive_edit(p, t[s:e], GLOBAL) == d
