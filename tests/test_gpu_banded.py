"""`sf_myers_banded` (the caller's re-capping, crates/map/src/lib.rs:851-911) on the device against
the oracle's restatement, for hundreds of reads in one call: the inputs of the reference's own
property test (lib.rs:3064-3120), all three modes, caps 0 to near u32::MAX, repeats, `distinct`,
packed input; and the reference's use as the tie-break makes it (a read against isoform texts)."""
import random

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A

from test_oracle_banded import case

pytestmark = pytest.mark.gpu


def groups(rng, count):
    reads, texts, caps = [], [], []
    for _ in range(count):
        r, ts, cap = case(rng)
        reads.append(r)
        texts.append(ts)
        caps.append(cap)
    return reads, texts, caps


@pytest.mark.parametrize("mode", [A.Mode.Infix, A.Mode.Prefix, A.Mode.Global])
def test_banded_batch_equals_oracle(mode):
    rng = random.Random(50 + int(mode))
    for cap in (0, 3, 16, 2**32 - 2):
        reads, texts, _ = groups(rng, 120)
        P, T, off = [], [], [0]
        for r, ts in zip(reads, texts):
            P += [r] * len(ts)
            T += ts
            off.append(len(T))
        ps = A.PairSet.from_seqs(P, T)
        got = A.banded_batch(ps, off, mode, cap)
        want = sum((O.banded(r, ts, int(mode), cap) for r, ts in zip(reads, texts)), [])
        assert [int(x) for x in got] == want, cap
        assert want == sum((O.sequential(r, ts, int(mode), cap) for r, ts in zip(reads, texts)), [])


def test_banded_distinct_and_single_read():
    rng = random.Random(3)
    for _ in range(40):
        read, texts, cap = case(rng)
        uniq = []
        for t in texts:
            if t not in uniq:
                uniq.append(t)
        assert A.banded(read, texts, A.Mode.Infix, cap) == O.banded(read, texts, O.INFIX, cap)
        assert A.banded(read, uniq, A.Mode.Infix, cap, distinct=True) == O.sequential(read, uniq, O.INFIX, cap)


def test_banded_long_reads_against_isoforms():
    # Reads of 500-1,500 bases against 3-12 texts each: the read's region with a few percent of
    # errors (isoforms), shifted windows of it, unrelated texts and repeats; cap 16 as the presets
    # use (lib.rs:231).  A group's later texts run at its first text's distance + cap over
    # Ukkonen's band, on every lane class up to 24 blocks.
    rng = random.Random(77)

    def mutate(s, rate):
        out = []
        for c in s:
            r = rng.random()
            if r < rate / 2:
                out.append(rng.randrange(4))
            elif r < 3 * rate / 4:
                out.extend([c, rng.randrange(4)])
            elif r >= rate:
                out.append(c)
        return out

    for mode in (A.Mode.Infix, A.Mode.Prefix, A.Mode.Global):
        reads, texts = [], []
        for _ in range(40):
            base = [rng.randrange(4) for _ in range(rng.randrange(500, 1500))]
            read = mutate(base, 0.06)
            ts = []
            for _ in range(rng.randrange(3, 12)):
                kind = rng.random()
                if kind < 0.6:
                    ts.append(mutate(base, rng.choice([0.01, 0.04, 0.10])))
                elif kind < 0.8:
                    lo = rng.randrange(0, 200)
                    ts.append(mutate(base[lo:], 0.03))
                elif kind < 0.9:
                    ts.append([rng.randrange(4) for _ in range(len(base))])
                elif ts:
                    ts.append(list(ts[rng.randrange(len(ts))]))
                else:
                    ts.append(list(base))
            reads.append(read)
            texts.append(ts)
        P, T, off = [], [], [0]
        for r, ts in zip(reads, texts):
            P += [r] * len(ts)
            T += ts
            off.append(len(T))
        ps = A.PairSet.from_seqs(P, T)
        for cap in (16, 0):
            got = A.banded_batch(ps, off, mode, cap)
            want = sum((O.sequential(r, ts, int(mode), cap) for r, ts in zip(reads, texts)), [])
            assert [int(x) for x in got] == want, (mode, cap)
