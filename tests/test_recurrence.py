"""The recurrence change the 16-bit packed kernel relies on, checked by brute force on small inputs.

That kernel opens a vertical gap (F) only from X = max(diagonal, F), H before the row's
horizontal gaps, not from H itself: a horizontal gap never turns straight into a vertical one.
For mis = 2 ext every H is unchanged (the turn ties with cutting its corner through one diagonal
step); for mis > 2 ext it is not, which the last test shows, so the condition is what carries it.
Plain Gotoh with the reference's gap cost open + ext * length (align.rs:1030-1060 in the banded
form), on a band of diagonals lo <= j - i <= hi, Global start.
"""
import random

NEG = -(10**9)


def gotoh_h(p, t, sc, lo, hi, f_from_x):
    mat, mis, go, ge = sc
    m, n = len(p), len(t)
    H = [[NEG] * (n + 1) for _ in range(m + 1)]
    E = [[NEG] * (n + 1) for _ in range(m + 1)]
    F = [[NEG] * (n + 1) for _ in range(m + 1)]
    X = [[NEG] * (n + 1) for _ in range(m + 1)]
    H[0][0] = X[0][0] = 0
    for j in range(1, n + 1):
        if lo <= j <= hi:
            H[0][j] = X[0][j] = -(go + ge * j)
    for i in range(1, m + 1):
        if lo <= -i <= hi:
            H[i][0] = X[i][0] = F[i][0] = -(go + ge * i)
        for j in range(1, n + 1):
            if not lo <= j - i <= hi:
                continue
            up = X[i - 1][j] if (f_from_x and i > 1) else H[i - 1][j]
            F[i][j] = max(F[i - 1][j], up - go) - ge
            E[i][j] = max(E[i][j - 1], X[i][j - 1] - go) - ge
            d = H[i - 1][j - 1] + (mat if p[i - 1] == t[j - 1] else -mis)
            X[i][j] = max(d, F[i][j])
            H[i][j] = max(X[i][j], E[i][j])
    return H


def pairs(rng, count):
    for _ in range(count):
        n = rng.randrange(0, 11)
        t = [rng.randrange(4) for _ in range(n)]
        p = list(t)
        for _ in range(rng.randrange(0, 3)):
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(0, 4))] + p[k + rng.randrange(0, 4):]
        lo, hi = -rng.randrange(0, 12), rng.randrange(0, 12)
        yield p, t, lo, hi


def test_f_from_x_keeps_every_h_when_mis_is_twice_ext():
    rng = random.Random(5)
    for sc in ((2, 4, 4, 2), (0, 2, 0, 1), (1, 4, 0, 2), (3, 2, 7, 1), (5, 6, 1, 3)):
        for p, t, lo, hi in pairs(rng, 400):
            # The first row's H is a horizontal gap from the origin; the kernel keeps it as the
            # up-value of row 1, so the comparison starts there too.
            assert gotoh_h(p, t, sc, lo, hi, True) == gotoh_h(p, t, sc, lo, hi, False), (sc, p, t, lo, hi)


def test_f_from_x_changes_h_when_mis_exceeds_twice_ext():
    rng = random.Random(7)
    sc = (1, 9, 1, 1)  # a mismatch costs more than a deletion next to an insertion
    differ = sum(gotoh_h(p, t, sc, lo, hi, True) != gotoh_h(p, t, sc, lo, hi, False)
                 for p, t, lo, hi in pairs(rng, 400))
    assert differ > 0
