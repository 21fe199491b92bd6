"""Calls from several host threads at once (the C ABI's thread-safety convention, sf_align.h:19):
one-shot Myers, affine score, CIGAR and clip/cut passes, and resident batches, eight threads side
by side, each result equal to the same call made alone, and a sample of them to the oracle.

ctypes releases the GIL for the length of every foreign call, so the threads' library calls
overlap on the host and their kernels on the device.
"""
import random
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu

THREADS = 8


def mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 2:
            out.append(rng.randrange(4))
        elif r < 3 * rate / 4:
            out.extend([c, rng.randrange(4)])
        elif r < rate:
            pass
        else:
            out.append(c)
    return out


def pair_set(seed):
    # Mixed lengths: short, workload-like, and patterns past one lane's 24 blocks (the pipelines).
    rng = random.Random(seed)
    P, T = [], []
    for i in range(160):
        m = rng.choice([0, 1, 40, 200, 700, 1000, 1600, 3000]) if i % 8 == 0 else rng.randrange(20, 1200)
        p = [rng.randrange(4) for _ in range(m)]
        t = [rng.randrange(4) for _ in range(rng.randrange(0, 60))] + mutate(rng, p, 0.08) + \
            [rng.randrange(4) for _ in range(rng.randrange(0, 60))]
        P.append(p)
        T.append(t)
    return A.PairSet.from_seqs(P, T)


def calls(ps, seed):
    """Every kind of call on one pair set; returns their outputs by name."""
    rng = random.Random(seed)
    band = rng.choice([16, 64, 128])
    geom = [len(ps.text(i)) + rng.randrange(0, 40) for i in range(len(ps))]
    cut = [rng.randrange(0, len(ps.text(i)) + 2) for i in range(len(ps))]
    out = {}
    out["infix_dist"], out["infix_end"], out["infix_start"] = A.myers_batch(ps, A.Mode.Infix, A.U32_MAX, locate=2)
    out["prefix_dist"], out["prefix_end"], _ = A.myers_batch(ps, A.Mode.Prefix, 40, locate=1)
    r = A.affine_batch(ps, A.Mode.Global, A.Affine.ONT, band)
    out["score"], out["score_col"] = r["score"], r["score_col"]
    rc = A.affine_batch(ps, A.Mode.Global, A.Affine.ONT, band, cigar=True)
    out["cigar_score"], out["cigar"], out["cigar_off"] = rc["score"], rc["cigar"], rc["cigar_off"]
    out["dp"] = np.asarray(A.affine_dp_batch(ps, A.Mode.Prefix, A.Affine.ONT, band, geom, (50, True), cut))
    b = A.Batch(ps)
    try:
        b.myers(A.Mode.Global, locate=0)
        b.affine(A.Mode.Infix, A.Affine.HIFI, band=band)
        b.sync()
        out["batch_dist"] = b.myers_results(0)[0]
        out["batch_score"] = b.affine_results(cigar=False)["score"]
    finally:
        b.close()
    return band, out


def test_eight_threads_equal_each_call_alone():
    sets = [pair_set(100 + t) for t in range(THREADS)]
    alone = [calls(ps, t) for t, ps in enumerate(sets)]
    with ThreadPoolExecutor(THREADS) as ex:
        for rnd in range(3):
            # Each round starts the threads in another order, so different calls overlap.
            order = list(range(THREADS))
            random.Random(rnd).shuffle(order)
            futs = {t: ex.submit(calls, sets[t], t) for t in order}
            for t in range(THREADS):
                band, got = futs[t].result(timeout=240)
                want_band, want = alone[t]
                assert band == want_band
                assert got.keys() == want.keys()
                for k in want:
                    assert np.array_equal(got[k], want[k]), (rnd, t, k)
    # The calls alone against the oracle, on a sample of each set.
    for t, ps in enumerate(sets):
        band, out = alone[t]
        dist, end, start, score = out["infix_dist"], out["infix_end"], out["infix_start"], out["score"]
        assert np.array_equal(out["cigar_score"], score)
        for i in range(0, len(ps), 23):
            p, x = ps.pattern(i), ps.text(i)
            d, e, s = O.locate(p, x, O.INFIX)
            assert (int(dist[i]), int(end[i]), int(start[i])) == tuple(A.U32_MAX if v is None else v for v in (d, e, s))
            assert int(score[i]) == O.affine_dp(p, x, O.GLOBAL, O.ONT, band).score


def test_threads_share_one_workload_batch_each():
    # A resident batch per thread over the same workload's pairs, run several times while the
    # others run: results equal a lone run.
    ps, bands = W.generate(1, 2048, 11)
    ref = A.Batch(ps)
    try:
        ref.myers(A.Mode.Infix, locate=1)
        ref.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=True)
        ref.sync()
        want_d = ref.myers_results(0)
        want_a = ref.affine_results(cigar=True)
    finally:
        ref.close()

    def run(_):
        b = A.Batch(ps)
        try:
            res = []
            for _ in range(3):
                b.myers(A.Mode.Infix, locate=1)
                b.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=True)
                b.sync()
                res.append((b.myers_results(0), b.affine_results(cigar=True)))
            return res
        finally:
            b.close()

    with ThreadPoolExecutor(THREADS) as ex:
        for res in ex.map(run, range(THREADS), timeout=240):
            for d, a in res:
                assert all(np.array_equal(x, y) for x, y in zip(d, want_d))
                for key in ("score", "score_col", "cigar", "cigar_off"):
                    assert np.array_equal(a[key], want_a[key]), key
