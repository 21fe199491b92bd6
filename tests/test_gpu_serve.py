"""The serving loop on the device: each read's results are the same alone and in the mix (and
equal the oracle's), oversized requests are refused by the HBM budget, and a failed batch fails
only its own reads (the failure is injected from the host; no kernel is made to fault)."""
import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import _native as N
from spearfish_amd import serve as S

pytestmark = pytest.mark.gpu

BAND = 64


@pytest.fixture(scope="module")
def workload():
    return S.generate_serve_workload(300, seed=7)


def store_codes(store):
    w = store.words
    k = np.arange(store.bases, dtype=np.int64)
    return ((w[k >> 4] >> (2 * (k & 15)).astype(np.uint32)) & 3).astype(np.uint8)


def test_results_alone_equal_in_mix_and_oracle(workload):
    store, reads = workload
    n = len(reads)
    alone, mix = S.Results(reads), S.Results(reads)
    reqs_alone = S.requests(reads, alone)
    with S.Server(store, band=BAND) as s:
        for i in range(n):
            t = s.submit(reqs_alone[i])
            assert s.wait(t) == N.SF_OK
        st = s.stats(reset=True)
        assert st.reads_done == n and st.batches == n
        reqs_mix = S.requests(reads, mix)
        s.replay(reqs_mix, S.poisson_arrivals(n, 200000.0, seed=1))
        st = s.stats()
        assert st.reads_done == n and st.batches < n  # reads shared batches
    for a in ("dist", "end", "score", "score_col"):
        assert np.array_equal(getattr(alone, a), getattr(mix, a)), a
    codes = store_codes(store)
    for i in range(0, n, 7):
        s0 = int(reads.seg0[i])
        r = reads.codes[int(reads.read_off[i]):int(reads.read_off[i]) + int(reads.read_len[i])]
        for k in range(int(reads.nseg[i])):
            q = s0 + k
            p = r[int(reads.seg_read_off[q]):int(reads.seg_read_off[q]) + int(reads.seg_read_len[q])]
            t = codes[int(reads.seg_tx_off[q]):int(reads.seg_tx_off[q]) + int(reads.seg_tx_len[q])]
            d, e, _ = O.locate(p, t, O.INFIX)
            assert (int(mix.dist[q]), int(mix.end[q])) == (d, e)
            dp = O.affine_dp(p, t, O.GLOBAL, O.ONT, BAND)
            assert (int(mix.score[q]), int(mix.score_col[q])) == (dp.score, dp.score_col)


def test_admission_refuses_requests_over_the_budget(workload):
    store, reads = workload
    res = S.Results(reads)
    reqs = S.requests(reads, res)
    big = int(np.argmax(reads.read_len))
    # A share of 64 KiB per slot holds 128 K bases and 256 segments: the longest read still fits,
    # a read whose bases exceed it is refused.
    with S.Server(store, hbm_budget_bytes=1 << 17, slots=2, band=BAND) as s:
        t = s.submit(reqs[big])
        assert s.wait(t) == N.SF_OK
        huge = np.zeros(200_000, np.uint8)
        r = N.ReadRequest(huge.ctypes.data, len(huge), 0, None, None, None, None, None, None, None, None)
        assert s.submit_status(r) == N.SF_ENOMEM


def test_injected_batch_failure_fails_only_its_reads(workload):
    store, reads = workload
    res = S.Results(reads)
    reqs = S.requests(reads, res, idx=range(6))
    with S.Server(store, slots=1, max_reads_per_batch=1, band=BAND, fail_batch=3) as s:
        status = []
        for i in range(6):
            status.append(s.wait(s.submit(reqs[i])))
        assert status == [N.SF_OK, N.SF_OK, N.SF_EDEVICE, N.SF_OK, N.SF_OK, N.SF_OK]
        st = s.stats()
        assert st.reads_done == 5 and st.reads_failed == 1


def block_count_reads(store, seed):
    """Reads whose segments have patterns of every block count from 1 to 33 (and the lengths at
    the class edges), each against a window of the store holding a mutated copy of it."""
    rng = np.random.default_rng(seed)
    codes = store_codes(store)
    lens = sorted({1, 63, 65, 255, 257, 319, 320, 321, 1535, 1537, 2047, 2048, 2049}
                  | {64 * b for b in range(1, 34)})
    rc, roff, rlen, nseg, sro, srl, sto, stl = [], [], [], [], [], [], [], []
    pos = 0
    for i in range(0, len(lens), 4):
        group = lens[i:i + 4]
        read = []
        for m in group:
            n = m + int(rng.integers(0, 60))
            t0 = int(rng.integers(0, store.bases - n))
            window = codes[t0:t0 + n]
            a = int(rng.integers(0, n - m + 1))
            p = window[a:a + m].copy()
            flip = rng.random(m) < 0.06
            p[flip] = (p[flip] + rng.integers(1, 4, int(flip.sum()))) % 4
            sro.append(len(read))
            srl.append(m)
            sto.append(t0)
            stl.append(n)
            read.extend(p.tolist())
        rc.extend(read)
        roff.append(pos)
        rlen.append(len(read))
        nseg.append(len(group))
        pos += len(read)
    nseg = np.array(nseg, np.uint32)
    seg0 = np.zeros(len(nseg), np.uint64)
    seg0[1:] = np.cumsum(nseg[:-1], dtype=np.uint64)
    return S.Reads(np.array(rc, np.uint8), np.array(roff, np.uint64), np.array(rlen, np.uint32), nseg, seg0,
                   np.array(sro, np.uint32), np.array(srl, np.uint32), np.array(sto, np.uint64),
                   np.array(stl, np.uint32))


@pytest.mark.parametrize("mode", [S.Mode.Infix, S.Mode.Prefix, S.Mode.Global])
def test_myers_every_block_count_alone_and_together(workload, mode):
    # A small batch runs patterns of 5-16 blocks on the 16-lane pipeline at one block a lane and
    # 17-24 at two (blocks in the last lanes, the leading lanes passing the top row through), 1-4
    # on the lane kernel, and every affine class at 64 lanes in one mixed launch beside them:
    # every segment's distance and end, score and score column equal the oracle's, served a read
    # at a time and all together.
    store, _ = workload
    reads = block_count_reads(store, seed=5 + int(mode))
    n = len(reads)
    alone, mix = S.Results(reads), S.Results(reads)
    with S.Server(store, band=BAND, myers_mode=mode) as s:
        ra = S.requests(reads, alone)
        for i in range(n):
            assert s.wait(s.submit(ra[i])) == N.SF_OK
        s.replay(S.requests(reads, mix), np.zeros(n))
    for a in ("dist", "end", "score", "score_col"):
        assert np.array_equal(getattr(alone, a), getattr(mix, a)), a
    codes = store_codes(store)
    for i in range(n):
        s0 = int(reads.seg0[i])
        r = reads.codes[int(reads.read_off[i]):int(reads.read_off[i]) + int(reads.read_len[i])]
        for k in range(int(reads.nseg[i])):
            q = s0 + k
            p = r[int(reads.seg_read_off[q]):int(reads.seg_read_off[q]) + int(reads.seg_read_len[q])]
            t = codes[int(reads.seg_tx_off[q]):int(reads.seg_tx_off[q]) + int(reads.seg_tx_len[q])]
            d, e, _ = O.locate(p, t, int(mode))
            assert (int(alone.dist[q]), int(alone.end[q])) == (d, e), (len(p), len(t))
            dp = O.affine_dp(p, t, O.GLOBAL, O.ONT, BAND)
            assert (int(alone.score[q]), int(alone.score_col[q])) == (dp.score, dp.score_col), (len(p), len(t))


def test_large_batches_equal_reads_alone():
    # Every read offered at once to one slot: batches grow past the small-batch limit (2,048
    # segments: one launch per class, pair records built on the device), and each read's results
    # equal those it gets served alone (pair records written on the host, one combined launch).
    store, reads = S.generate_serve_workload(4000, seed=11)
    n = len(reads)
    alone, mix = S.Results(reads), S.Results(reads)
    with S.Server(store, slots=1, band=BAND) as s:
        ra = S.requests(reads, alone)
        for i in range(n):
            assert s.wait(s.submit(ra[i])) == N.SF_OK
        s.stats(reset=True)
        s.replay(S.requests(reads, mix), np.zeros(n))
        st = s.stats()
        assert st.reads_done == n and st.mean_reads_per_batch * 8 > 2048, st.mean_reads_per_batch
    for a in ("dist", "end", "score", "score_col"):
        assert np.array_equal(getattr(alone, a), getattr(mix, a)), a
    codes = store_codes(store)
    for i in range(0, n, 97):
        s0 = int(reads.seg0[i])
        r = reads.codes[int(reads.read_off[i]):int(reads.read_off[i]) + int(reads.read_len[i])]
        for k in range(int(reads.nseg[i])):
            q = s0 + k
            p = r[int(reads.seg_read_off[q]):int(reads.seg_read_off[q]) + int(reads.seg_read_len[q])]
            t = codes[int(reads.seg_tx_off[q]):int(reads.seg_tx_off[q]) + int(reads.seg_tx_len[q])]
            d, e, _ = O.locate(p, t, O.INFIX)
            assert (int(mix.dist[q]), int(mix.end[q])) == (d, e)
            dp = O.affine_dp(p, t, O.GLOBAL, O.ONT, BAND)
            assert (int(mix.score[q]), int(mix.score_col[q])) == (dp.score, dp.score_col)
