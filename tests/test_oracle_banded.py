"""The oracle's restatement of the caller's re-capping, `Mapper::banded`
(crates/map/src/lib.rs:851-911: distinct texts in batches of eight at the threshold in force at
each batch's first text, re-capped in order), against its sequential definition (lib.rs:3016-3034:
one distance_with per distinct text in order), on inputs shaped as the reference's own property
test shapes them (lib.rs:3064-3120): mutated excerpts of the read's source, unrelated and empty
texts, repeats, caps from 0 to near u32::MAX; with and without `distinct`."""
import random

import pytest

from oracle import sf_oracle as O

U32 = 2**32 - 1


def mutate(s, edits):
    v = list(s)
    for pos, b, kind in edits:
        if not v:
            if kind == 1:
                v.append(b)
            continue
        i = pos % len(v)
        if kind == 0:
            v[i] = b
        elif kind == 1:
            v.insert(i, b)
        else:
            del v[i]
    return v


def case(rng):
    base = [rng.randrange(4) for _ in range(rng.randrange(0, 300))]
    edits = lambda k: [(rng.randrange(300), rng.randrange(4), rng.randrange(3)) for _ in range(rng.randrange(k))]
    read = mutate(base, edits(30))
    texts = []
    for _ in range(rng.randrange(1, 24)):
        if rng.random() < 0.2:
            texts.append([e[1] for e in edits(40)])
        else:
            t = mutate(base, edits(40))
            lo = min(rng.randrange(40), len(t))
            hi = max(len(t) - rng.randrange(40), lo)
            texts.append(t[lo:hi])
    for _ in range(rng.randrange(8)):
        t = texts[rng.randrange(len(texts))]
        texts.insert(rng.randrange(len(texts) + 1), list(t))
    cap = rng.choice([rng.randrange(6), rng.randrange(6, 40), U32 - rng.randrange(4)])
    return read, texts, cap


@pytest.mark.parametrize("mode", [O.INFIX, O.PREFIX, O.GLOBAL])
def test_banded_equals_sequential(mode):
    rng = random.Random(1000 + mode)
    for _ in range(150):
        read, texts, cap = case(rng)
        want = O.sequential(read, texts, mode, cap)
        assert O.banded(read, texts, mode, cap) == want
        # `distinct` on the texts without their repeats gives those texts' values.
        uniq, idx = [], []
        for i, t in enumerate(texts):
            if t not in uniq:
                uniq.append(t)
                idx.append(i)
        assert O.banded(read, uniq, mode, cap, distinct=True) == O.sequential(read, uniq, mode, cap)


def test_sequential_is_the_capped_exact_distance():
    # The definition itself, restated from its parts: k = best so far + cap (saturating), the exact
    # distance where it is <= k, else k + 1; a repeat is its first occurrence's value re-capped.
    rng = random.Random(7)
    for _ in range(100):
        read, texts, cap = case(rng)
        best, seen, want = U32, {}, []
        for t in texts:
            k = min(U32, best + cap)
            key = tuple(t)
            if key in seen:
                d = seen[key]
                want.append(d if d <= k else min(U32, k + 1))
                continue
            e = O.naive_edit(read, t, O.INFIX)
            d = e if e <= k else min(U32, k + 1)
            best = min(best, d)
            seen[key] = d
            want.append(d)
        assert O.sequential(read, texts, O.INFIX, cap) == want
