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


def test_injected_batch_failure_fails_only_its_reads(workload, monkeypatch):
    # SF_SERVE_FAIL_BATCH (read when a server is created): the n-th batch fails without a launch.
    store, reads = workload
    res = S.Results(reads)
    reqs = S.requests(reads, res, idx=range(6))
    monkeypatch.setenv("SF_SERVE_FAIL_BATCH", "3")
    with S.Server(store, slots=1, max_reads_per_batch=1, band=BAND) as s:
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
    # Every read offered at once to one slot that waits to fill its batch (up to 0.5 s after the
    # first read): batches grow past the small-batch limit (2,048 segments: one launch per class,
    # pair records built on the device), and each read's results equal those it gets served alone
    # (pair records written on the host, one combined launch).
    store, reads = S.generate_serve_workload(4000, seed=11)
    n = len(reads)
    alone, mix = S.Results(reads), S.Results(reads)
    with S.Server(store, slots=1, band=BAND) as s:
        ra = S.requests(reads, alone)
        for i in range(n):
            assert s.wait(s.submit(ra[i])) == N.SF_OK
    with S.Server(store, slots=1, band=BAND, batch_wait_us=500_000) as s:
        s.replay(S.requests(reads, mix), np.zeros(n))
        st = s.stats()
        assert st.reads_done == n and st.max_segments_per_batch > 2048, st.max_segments_per_batch
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


def test_timed_out_wait_then_wait_again_and_cancel(workload):
    # One slot, one read a batch, 120 reads offered at once: the last ones are still queued right
    # after submission.  A wait that times out leaves its request queued and its ticket valid (the
    # next wait returns the result); a cancelled request completes with SF_ECANCELED and its
    # ticket is retired by that wait.
    store, reads = workload
    n = 120
    alone, res = S.Results(reads), S.Results(reads)
    ra, rr = S.requests(reads, alone, idx=range(n)), S.requests(reads, res, idx=range(n))
    with S.Server(store, slots=1, max_reads_per_batch=1, band=BAND) as s:
        for i in range(n):
            assert s.wait(s.submit(ra[i])) == N.SF_OK
        tickets = [s.submit(rr[i]) for i in range(n)]
        assert s.wait(tickets[-2], 0.0) == N.SF_ETIMEDOUT
        assert s.wait(tickets[-1], 0.0) == N.SF_ETIMEDOUT
        assert s.cancel(tickets[-1]) == N.SF_OK
        assert s.wait(tickets[-1]) == N.SF_ECANCELED
        assert s.wait(tickets[-1]) == N.SF_EINVAL  # retired
        assert s.cancel(tickets[-1]) == N.SF_EINVAL
        for t in tickets[:-1]:
            assert s.wait(t) == N.SF_OK
    last = reads.seg0[n - 1]
    for a in ("dist", "end", "score", "score_col"):
        assert np.array_equal(getattr(alone, a)[:last], getattr(res, a)[:last]), a
    # The cancelled read was still queued: nothing was written to its outputs.
    assert (res.dist[last:last + reads.nseg[n - 1]] == 0xFFFFFFFF).all()


def test_cancel_many_and_concurrent_waiters(workload):
    # One slot, one read a batch, 200 reads offered at once: the last read waits behind 199
    # batches.  Two threads wait on its ticket (both get its status, then the ticket is retired)
    # while the main thread cancels the reads before it in the second half.
    import threading
    store, reads = workload
    n = 200
    alone, res = S.Results(reads), S.Results(reads)
    ra, rr = S.requests(reads, alone, idx=range(n)), S.requests(reads, res, idx=range(n))
    with S.Server(store, slots=1, max_reads_per_batch=1, band=BAND) as s:
        for i in range(n):
            assert s.wait(s.submit(ra[i])) == N.SF_OK
        s.stats(reset=True)
        tickets = [s.submit(rr[i]) for i in range(n)]
        assert s.wait(tickets[-1], 0.0) == N.SF_ETIMEDOUT
        got = []
        th = [threading.Thread(target=lambda: got.append(s.wait(tickets[-1]))) for _ in range(2)]
        for t in th:
            t.start()
        cancelled = [i for i in range(n // 2, n - 1) if s.cancel(tickets[i]) == N.SF_OK]
        for t in th:
            t.join()
        assert got == [N.SF_OK, N.SF_OK]
        assert s.wait(tickets[-1]) == N.SF_EINVAL
        assert len(cancelled) > 0
        for i in range(n - 1):
            st = s.wait(tickets[i])
            assert st == (N.SF_ECANCELED if i in cancelled else N.SF_OK), (i, st)
        assert s.stats().reads_done == n - len(cancelled)
    for i in range(n):
        if i in cancelled:
            continue
        a, b = int(reads.seg0[i]), int(reads.seg0[i]) + int(reads.nseg[i])
        for f in ("dist", "end", "score", "score_col"):
            assert np.array_equal(getattr(alone, f)[a:b], getattr(res, f)[a:b]), (i, f)
