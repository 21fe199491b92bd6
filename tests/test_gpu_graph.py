"""Persistent mixed launches and captured steps on the uneven workload (many shape classes, so the
Myers and score runs take the mixed launches): the results of the persistent schedule, and of every
replay of a captured step, equal the eager one-workgroup-a-block runs, which equal the oracle.

Replays run several times with eager work and synchronisations between them, each checked."""
import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import _native as N
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven():
    ps, bands = W.generate(3, 40000, 2)
    return ps, bands


def step(b, bands, mode=A.Mode.Infix):
    b.myers(mode, locate=1)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)


def results(b):
    d, e, _ = b.myers_results(1)
    r = b.affine_results()
    return d, e, r["score"], r["score_col"]


def same(x, y):
    return all(np.array_equal(p, q) for p, q in zip(x, y))


@pytest.fixture
def persistent(request):
    lib = N.load()
    yield lib
    N.check(lib.sf_set_persistent(0))


def test_persistent_and_replays_equal_eager(uneven, persistent):
    ps, bands = uneven
    lib = persistent
    b = A.Batch(ps)
    N.check(lib.sf_set_persistent(0))
    step(b, bands)
    eager = results(b)
    for i in range(0, len(ps), 211):
        p, t = ps.pattern(i), ps.text(i)
        assert (int(eager[0][i]), int(eager[1][i])) == O.locate(p, t, O.INFIX)[:2], i
        dp = O.affine_dp(p, t, O.GLOBAL, O.ONT, int(bands[i]))
        assert (int(eager[2][i]), int(eager[3][i])) == (dp.score, dp.score_col), i
    # Persistent launches, run twice in a row (the queues reset themselves between launches).
    for pers in (1, 2, 3):
        N.check(lib.sf_set_persistent(pers))
        for _ in range(2):
            step(b, bands)
            assert same(results(b), eager), pers
    # Captured steps: eager schedule and persistent schedule, replayed with eager work between.
    for pers in (0, 3):
        N.check(lib.sf_set_persistent(pers))
        b.capture_begin()
        step(b, bands)
        b.capture_end()
        for rep in range(3):
            b.myers(A.Mode.Prefix, locate=1)  # eager work on the same outputs, then the replay
            b.sync()
            b.replay(1 + rep)
            assert same(results(b), eager), (pers, rep)
    t = b.timing()
    assert t.graph_replays == 2 * (1 + 2 + 3)
    b.close()


def test_capture_refuses_blocking_calls(uneven):
    ps, bands = uneven
    sub = A.PairSet(ps.pat, ps.pat_off[:500], ps.pat_len[:500], ps.txt, ps.txt_off[:500], ps.txt_len[:500])
    b = A.Batch(sub)
    lib = N.load()
    b.myers(A.Mode.Infix, locate=1)
    b.sync()
    # A CIGAR run, or a score run without its launch plan, cannot be captured; the capture fails.
    b.capture_begin()
    b.myers(A.Mode.Infix, locate=1)
    assert lib.sf_batch_affine(b.h, 2, A.Affine.ONT.c(), A._bands(64, 500).ctypes.data, 0) == N.SF_EINVAL
    assert lib.sf_batch_capture_end(b.h) == N.SF_EINVAL
    assert lib.sf_batch_replay(b.h, 1) == N.SF_EINVAL  # nothing captured
    # The batch still works eagerly afterwards.
    b.affine(A.Mode.Global, A.Affine.ONT, band=64, cigar=True)
    r = b.affine_results(cigar=True)
    s, col, ts, cig = O.align(sub.pattern(7), sub.text(7), O.GLOBAL, O.ONT, 64)
    assert int(r["score"][7]) == s
    b.close()
