"""GPU parity on the uniform workload at its full size: 1,000,000 pairs of 1,000 bp targets with
queries at 6% error, band 128, exactly as the benchmark runs it.  At this size the launch planner
picks the production shapes: the 4-lane packed affine score kernels (sixteen pairs a wave), the
16-lane packed traceback kernel, and the Myers lane kernel's full-wave path for 16-block patterns,
launched alone.  Every pair's distance and
affine score are compared with the oracle on all host threads; ends, start columns and CIGARs on a
strided sample.
"""
import os

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu

N_PAIRS = 1_000_000
THREADS = os.cpu_count() or 1
STRIDE = 4999


@pytest.fixture(scope="module")
def uniform():
    ps, bands = W.generate(2, N_PAIRS, 0, threads=min(THREADS, 16))
    assert np.all(bands == 128)
    b = A.Batch(ps)
    yield ps, b
    b.close()


def oracle_dist(ps, mode, k):
    return O.batch_myers(ps.pat, ps.pat_off, ps.pat_len, ps.txt, ps.txt_off, ps.txt_len, int(mode), k, THREADS)


@pytest.mark.parametrize("mode", [A.Mode.Infix, A.Mode.Prefix, A.Mode.Global])
@pytest.mark.parametrize("k", [A.U32_MAX, 60], ids=["k=inf", "k=60"])
def test_myers_every_pair(uniform, mode, k):
    ps, b = uniform
    blocks = (ps.pat_len.astype(np.int64) + 63) // 64
    assert np.mean(blocks == 16) > 0.9  # the full-wave 16-block path carries the batch
    b.myers(mode, k, locate=1)
    dist, end, _ = b.myers_results(1)
    want = oracle_dist(ps, mode, k)
    bad = np.flatnonzero(dist != want)
    assert bad.size == 0, (bad[:10], dist[bad[:10]], want[bad[:10]])
    for i in range(0, N_PAIRS, STRIDE):
        d, e, _ = O.locate(ps.pattern(i), ps.text(i), int(mode), k)
        assert (int(dist[i]), int(end[i])) == (A.U32_MAX if d is None else d, A.U32_MAX if e is None else e), i


def test_affine_scores_every_pair_and_cigar_sample(uniform):
    ps, b = uniform
    b.affine(A.Mode.Global, A.Affine.ONT, band=128)
    r = b.affine_results()
    args = (ps.pat, ps.pat_off, ps.pat_len, ps.txt, ps.txt_off, ps.txt_len, O.GLOBAL, O.ONT, 128, THREADS)
    if O.avx512_available():
        score, col = O.batch_affine(*args, avx512=True)  # equal to the scalar rows (test_oracle_golden)
        idx = np.arange(N_PAIRS)
    else:
        idx = np.arange(0, N_PAIRS, 7)
        sub = A.PairSet(ps.pat, ps.pat_off[idx], ps.pat_len[idx], ps.txt, ps.txt_off[idx], ps.txt_len[idx])
        score, col = O.batch_affine(sub.pat, sub.pat_off, sub.pat_len, sub.txt, sub.txt_off, sub.txt_len,
                                    O.GLOBAL, O.ONT, 128, THREADS)
    assert np.array_equal(r["score"][idx], score)
    assert np.array_equal(r["score_col"][idx], col.astype(np.uint32))
    b.affine(A.Mode.Global, A.Affine.ONT, band=128, cigar=True)
    rc = b.affine_results(cigar=True)
    assert np.array_equal(rc["score"], r["score"])
    for i in range(0, N_PAIRS, STRIDE):
        s, c0, ts, cig = O.align(ps.pattern(i), ps.text(i), O.GLOBAL, O.ONT, 128)
        got = [int(x) for x in rc["cigar"][int(rc["cigar_off"][i]):int(rc["cigar_off"][i + 1])]]
        assert (int(rc["score"][i]), got) == (s, cig), i
