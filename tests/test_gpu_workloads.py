"""GPU parity on the uneven and long-segment workloads, end to end through device-resident batches.

Uneven (workload 3): log-normal lengths 100..20,000 bp, 5-15% error, per-pair band
max(32, 1.2|m - n| + 32) capped at 512; Myers Infix + affine Global with CIGAR.
Long (workload 4): queries 2-10 kb, targets 1.05x; Myers Infix distance, end and start, then
affine Global band 256 with CIGAR on the located window text[start..end).
"""
import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu


def cig(r, i):
    return [int(x) for x in r["cigar"][int(r["cigar_off"][i]):int(r["cigar_off"][i + 1])]]


def test_uneven_workload_subset():
    ps, bands = W.generate(3, 160, 0)
    b = A.Batch(ps)
    b.myers(A.Mode.Infix, locate=2)
    dist, end, start = b.myers_results(2)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=True)
    r = b.affine_results(cigar=True)
    for i in range(len(ps)):
        p, t = ps.pattern(i), ps.text(i)
        assert (int(dist[i]), int(end[i]), int(start[i])) == O.locate(p, t, O.INFIX), i
        s, col, ts, c = O.align(p, t, O.GLOBAL, O.ONT, int(bands[i]))
        assert (int(r["score"][i]), int(r["score_col"][i]), cig(r, i)) == (s, col, c), (i, len(p), len(t), int(bands[i]))
    b.close()


@pytest.mark.parametrize("budget_mb", [None, "16"], ids=["one-chunk", "16MB-chunks"])
def test_long_workload_located_windows(budget_mb, monkeypatch):
    # With 16 MB of traceback bits a chunk (SF_TB_BUDGET_MB), the pairs' bits (about 40 MB) run in
    # several double-buffered chunks, walks beside the next chunk's DP.
    if budget_mb:
        monkeypatch.setenv("SF_TB_BUDGET_MB", budget_mb)
    ps, bands = W.generate(4, 12, 0)
    b = A.Batch(ps)
    b.myers(A.Mode.Infix, locate=2)
    dist, end, start = b.myers_results(2)
    b.use_located(True)
    b.affine(A.Mode.Global, A.Affine.ONT, band=256, cigar=True)
    r = b.affine_results(cigar=True)
    for i in range(len(ps)):
        p, t = ps.pattern(i), ps.text(i)
        d, e, s = O.locate(p, t, O.INFIX)
        assert (int(dist[i]), int(end[i]), int(start[i])) == (d, e, s), i
        sc, col, ts, c = O.align(p, t[s:e], O.GLOBAL, O.ONT, 256)
        assert (int(r["score"][i]), cig(r, i)) == (sc, c), i
    # Restoring the texts gives the whole-text results again.
    b.use_located(False)
    b.myers(A.Mode.Infix, locate=2)
    assert np.array_equal(b.myers_results(2)[0], dist)
    b.close()


def test_uneven_steps_queued_back_to_back():
    # A Myers run leaves the batch's extra streams open and the score run after it starts beside
    # its last bins (both read only the sequences); the next Myers run, and the results, wait for
    # them.  Steps queued without a sync in between (as the bench runs them), a Prefix run queued
    # right after an Infix one, and a CIGAR run after a score run all give the results of runs
    # made one at a time, which match the oracle.
    ps, bands = W.generate(3, 3000, 4)
    b = A.Batch(ps)
    b.myers(A.Mode.Infix, locate=1)
    dist0, end0, _ = b.myers_results(1)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    r0 = b.affine_results()
    for _ in range(3):
        b.myers(A.Mode.Infix, locate=1)
        b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    dist1, end1, _ = b.myers_results(1)
    r1 = b.affine_results()
    assert np.array_equal(dist0, dist1) and np.array_equal(end0, end1)
    assert np.array_equal(r0["score"], r1["score"]) and np.array_equal(r0["score_col"], r1["score_col"])
    b.myers(A.Mode.Infix, locate=1)
    b.myers(A.Mode.Prefix, locate=1)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=True)
    dist2, end2, _ = b.myers_results(1)
    r2 = b.affine_results(cigar=True)
    assert np.array_equal(r0["score"], r2["score"])
    for i in range(0, len(ps), 37):
        p, t = ps.pattern(i), ps.text(i)
        assert (int(dist0[i]), int(end0[i])) == O.locate(p, t, O.INFIX)[:2], i
        assert (int(dist2[i]), int(end2[i])) == O.locate(p, t, O.PREFIX)[:2], i
        s, col, ts, c = O.align(p, t, O.GLOBAL, O.ONT, int(bands[i]))
        assert (int(r0["score"][i]), int(r0["score_col"][i])) == (s, col), i
        assert cig(r2, i) == c, i
    b.close()
