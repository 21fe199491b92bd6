"""Myers patterns longer than the largest register-resident class (1,536 blocks, 98,304 bases), run
as stripes of 1,536 blocks whose column deltas pass through scratch (myers_pipe_body<64, 24, true>),
against the oracle: distance, end and start in all modes, at k unbounded and finite, two and three
stripes, in one batch with short pairs."""
import random

import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A

pytestmark = pytest.mark.gpu


def long_pairs(seed):
    rng = random.Random(seed)
    P, T = [], []
    for m in (98305, 120000, 150001, 200000, 64, 3000):
        t = [rng.randrange(4) for _ in range(int(m * 1.05) + 50)]
        a = rng.randrange(0, len(t) - m + 1)
        p = list(t[a:a + m])
        for _ in range(m // 12):  # about 8% edits
            i = rng.randrange(len(p))
            k = rng.randrange(3)
            if k == 0:
                p[i] = rng.randrange(4)
            elif k == 1:
                p.insert(i, rng.randrange(4))
            else:
                del p[i]
        P.append(p)
        T.append(t)
    return P, T


@pytest.mark.parametrize("mode", [A.Mode.Infix, A.Mode.Prefix, A.Mode.Global])
def test_striped_patterns_equal_the_oracle(mode):
    P, T = long_pairs(40 + int(mode))
    ps = A.PairSet.from_seqs(P, T)
    for k in (A.U32_MAX, 20000):
        dist, end, start = A.myers_batch(ps, mode, k, locate=2)
        for i in range(len(ps)):
            d, e, s = O.locate(P[i], T[i], int(mode), k)
            want = tuple(A.U32_MAX if x is None else x for x in (d, e, s))
            assert (int(dist[i]), int(end[i]), int(start[i])) == want, (i, len(P[i]), k)
