"""GPU parity: the HIP kernels through the C ABI against the CPU oracle, bit for bit.

Inputs are seeded: the parity-set workload (1) and hand-made edge cases.  Every comparison is
exact (integer scores, distances, coordinates and CIGAR ops).
"""
import random

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu

MODES = (A.Mode.Infix, A.Mode.Prefix, A.Mode.Global)
U32 = A.U32_MAX


def opt(x):
    return U32 if x is None else x


def test_lane_moves():
    import ctypes as C
    from spearfish_amd import _native as N
    out = (C.c_int32 * 320)()
    N.check_diag(N.load_diag().sf_selftest_lanes(C.cast(out, C.c_void_p)))
    v = list(out)
    assert v[0:64] == [l + 1 for l in range(63)] + [-1], "wave_shl:1"
    assert v[64:128] == [-1] + list(range(63)), "wave_shr:1"
    assert v[128:192] == [(-1 if l % 16 == 0 else l - 1) for l in range(64)], "row_shr:1"
    vals = [l * 3 % 17 for l in range(64)]
    assert v[192:256] == [-100] + [max(vals[:l]) for l in range(1, 64)], "exclusive max scan"
    assert v[256:320] == [l ^ 5 for l in range(64)], "shfl_xor"


def edge_pairs():
    rng = random.Random(21)
    P, T = [], []
    for m in (0, 1, 2, 63, 64, 65, 127, 128, 129, 500, 1000, 1024, 1025, 2048):
        for n in (0, 1, 5, 64, 200, 1100):
            p = [rng.randrange(4) for _ in range(m)]
            t = [rng.randrange(4) for _ in range(n)]
            if n >= m and m > 0 and rng.random() < 0.5:
                s = rng.randrange(n - m + 1)
                t[s:s + m] = p
            P.append(p)
            T.append(t)
    # codes above 3 are masked (align.rs:78, 120)
    P.append([4, 5, 6, 7, 0, 1]), T.append([0, 1, 2, 3, 8, 9, 10])
    return P, T


def check_myers(ps, mode, k):
    dist, end, start = A.myers_batch(ps, mode, k, locate=2)
    for i in range(len(ps)):
        p, t = ps.pattern(i), ps.text(i)
        d, e, s = O.locate(p, t, int(mode), k)
        assert (int(dist[i]), int(end[i]), int(start[i])) == (opt(d), opt(e), opt(s)), (i, len(p), len(t), mode, k)


@pytest.mark.parametrize("mode", MODES)
def test_myers_parity_set(mode):
    ps, _ = W.generate(1, 1024, 0)
    check_myers(ps, mode, U32)
    check_myers(ps, mode, 40)


@pytest.mark.parametrize("mode", MODES)
def test_myers_edges(mode):
    P, T = edge_pairs()
    ps = A.PairSet.from_seqs(P, T)
    for k in (U32, 0, 3, 100):
        check_myers(ps, mode, k)


def test_myers_reference_examples():
    # align.rs:2908-2926 through the device
    assert A.infix_distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3]) == 0
    assert A.infix_distance([0, 1, 2, 3], [0, 1, 3]) == 1
    assert A.infix_distance([0, 1, 2, 3], []) == 4
    assert A.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], A.Mode.Prefix, 10) == 2
    assert A.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], A.Mode.Global, 10) == 3
    assert A.distance([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3], A.Mode.Global, 2) is None
    assert A.distance([], [0, 1], A.Mode.Global, 5) == 2
    pat = A.Pattern([0, 1, 2, 3])
    assert A.batch_distances(pat, [[3, 3, 0, 1, 2, 3, 3], [0, 1, 3], []], A.Mode.Infix, U32) == [0, 1, 4]


def check_affine(ps, mode, sc, bands, cigar=True):
    r = A.affine_batch(ps, mode, sc, bands, cigar=cigar)
    bands = np.broadcast_to(np.asarray(A.USIZE_MAX if bands is None else bands, np.uint64), (len(ps),))
    for i in range(len(ps)):
        p, t = ps.pattern(i), ps.text(i)
        b = int(bands[i])
        bo = None if b == A.USIZE_MAX else b
        dp = O.affine_dp(p, t, int(mode), (sc.mat, sc.mis, sc.open, sc.ext), bo)
        assert (int(r["score"][i]), int(r["score_col"][i])) == (dp.score, dp.score_col), (i, len(p), len(t), mode, b)
        if cigar:
            want = O.align(p, t, int(mode), (sc.mat, sc.mis, sc.open, sc.ext), bo)
            o0, o1 = int(r["cigar_off"][i]), int(r["cigar_off"][i + 1])
            got = [int(x) for x in r["cigar"][o0:o1]]
            if want is None:
                assert got == []
            else:
                assert (int(r["tstart"][i]), got) == (want[2], want[3]), (i, len(p), len(t), mode, b,
                                                                          O.cigar_str(got), O.cigar_str(want[3]))


def test_affine_parity_set_global_band64():
    ps, band = W.generate(1, 1024, 0)
    check_affine(ps, A.Mode.Global, A.Affine.ONT, band)


@pytest.fixture(params=[0, 1, 2], ids=["planned", "fewest-lanes-8", "fewest-lanes-4"])
def lane_policy(request):
    """Lanes per pair as the launch planner picks them (small batches: mostly 64, every pair's
    rows outlasting the rest), or every pair in its fewest-lanes shape (8 or 16; or 4 first),
    then back."""
    lib = A.N.load()
    A.N.check(lib.sf_set_affine_lp8(request.param))
    yield request.param
    A.N.check(lib.sf_set_affine_lp8(0))


def test_affine_traceback_in_many_chunks(monkeypatch):
    # Traceback bits in 16 MB chunks (SF_TB_BUDGET_MB, read at each call): the 2,048 pairs of the
    # parity-set generator (band 64) need about ten, so the double-buffered bits, each chunk's walk
    # beside the next chunk's DP and the largest bin's last piece run last all take part.
    monkeypatch.setenv("SF_TB_BUDGET_MB", "16")
    ps, band = W.generate(1, 2048, 3)
    check_affine(ps, A.Mode.Global, A.Affine.ONT, band)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sc", [A.Affine.ONT, A.Affine.HIFI, A.Affine(1, 1, 0, 1), A.Affine(3, 2, 5, 1)])
def test_affine_small_random(mode, sc, lane_policy):
    rng = random.Random(31 + int(mode))
    P, T, B = [], [], []
    for _ in range(300):
        n = rng.randrange(0, 300)
        t = [rng.randrange(4) for _ in range(n)]
        p = []
        for x in t:
            u = rng.random()
            if u < 0.05:
                p.append((x + 1) & 3)
            elif u < 0.08:
                p += [x, rng.randrange(4)]
            elif u < 0.11:
                pass
            else:
                p.append(x)
        if rng.random() < 0.2:
            p = p[rng.randrange(len(p) + 1):]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 1, 2, 5, 16, 64, 200, A.USIZE_MAX]))
    ps = A.PairSet.from_seqs(P, T)
    check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=mode != A.Mode.Infix or True)


@pytest.fixture
def lanes64():
    """Every band on 64 lanes per pair for the test, then the default back."""
    lib = A.N.load()
    lib.sf_set_affine_lp16_max(0)
    yield
    lib.sf_set_affine_lp16_max(512)


@pytest.mark.parametrize("mode", MODES)
def test_affine_64_lane_kernel(mode, lanes64):
    # The small bands that normally run 16 lanes per pair, through the 64-lane kernel.
    rng = random.Random(7 + int(mode))
    P, T, B = [], [], []
    for _ in range(200):
        n = rng.randrange(0, 260)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.08 else rng.randrange(4) for x in t]
        if rng.random() < 0.3:
            p = p[rng.randrange(len(p) + 1):]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 1, 5, 16, 64, 130, A.USIZE_MAX]))
    ps = A.PairSet.from_seqs(P, T)
    # mis = 2 ext: the 16-bit kernels on 64 lanes; (1, 1, 0, 1): the 32-bit ones.
    for sc in (A.Affine.ONT, A.Affine(3, 2, 5, 1), A.Affine(1, 1, 0, 1)):
        check_affine(ps, mode, sc, np.array(B, np.uint64))


@pytest.mark.parametrize("mode", MODES)
def test_affine_scores_16bit(mode, lane_policy):
    # Score-only runs with mis = 2 ext (ONT, HIFI) take the 16-bit kernels where their ranges hold:
    # random pairs, all-match and gap-heavy extremes, and lengths around their limits at band 16,
    # so that the 32-bit and both 16-bit kernels run in one batch.
    rng = random.Random(53 + int(mode))
    P, T, B = [], [], []
    nmax = 450 if mode == A.Mode.Infix else 1500  # Infix is unbanded: m + n + 1 slots a row
    for _ in range(300):
        n = rng.randrange(0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.1 else rng.randrange(4) for x in t]
        if rng.random() < 0.2:
            p = p[:rng.randrange(len(p) + 1)] + [rng.randrange(4) for _ in range(rng.randrange(40))]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 1, 7, 16, 64, 128, 200]))
    for m in ((3000, 500) if mode != A.Mode.Infix else (400, 300)):
        t = [rng.randrange(4) for _ in range(m)]
        P += [t, [(x + 1) & 3 for x in t], t[: 2 * m // 3]]
        T += [t, t, t]
        B += [128, 128, 64]
    if mode != A.Mode.Infix:
        # Around the three-term-max kernel's limit (values below 0x7C00, m near 5,200 for ONT at
        # band 16) and the 16-bit kernel's (m near 10,700).
        for m in (5150, 5250, 5350, 10600, 10700, 10800, 11000):
            t = [rng.randrange(4) for _ in range(m)]
            P.append([x if rng.random() > 0.05 else rng.randrange(4) for x in t])
            T.append(t)
            B.append(16)
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.ONT, A.Affine.HIFI):
        check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=False)


@pytest.fixture
def lanes8():
    """Score-only packed runs on 8 lanes per pair wherever that kernel's range holds, then the
    planner's choice back."""
    lib = A.N.load()
    A.N.check(lib.sf_set_affine_lp8(1))
    yield
    A.N.check(lib.sf_set_affine_lp8(0))


@pytest.mark.parametrize("mode", MODES)
def test_affine_scores_8_lane_kernel(mode, lanes8):
    # Every band width up to the 8-lane kernel's 288 slots (1 to 18 registers a lane), random and
    # gap-heavy pairs, lengths around its value limit (m near 5,200 for ONT at band 16).
    rng = random.Random(67 + int(mode))
    P, T, B = [], [], []
    nmax = 140 if mode == A.Mode.Infix else 900
    for _ in range(500):
        n = rng.randrange(0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.1 else rng.randrange(4) for x in t]
        for _ in range(rng.randrange(0, 4)):
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(0, 9))] + p[k + rng.randrange(0, 9):]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 1, 3, 7, 8, 15, 16, 40, 64, 100, 127, 128, 143]))
    if mode != A.Mode.Infix:
        for m in (5150, 5250, 5350):
            t = [rng.randrange(4) for _ in range(m)]
            P.append([x if rng.random() > 0.05 else rng.randrange(4) for x in t])
            T.append(t)
            B.append(16)
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.ONT, A.Affine.HIFI, A.Affine(0, 2, 0, 1), A.Affine(3, 2, 7, 1)):
        check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=False)


@pytest.fixture
def lanes4():
    """Score-only packed runs on 4 lanes per pair wherever that kernel's range holds (bands of 1
    to 288 slots: 36 to 72 slots a lane, 66 and 70 among them), then the planner's choice back."""
    lib = A.N.load()
    A.N.check(lib.sf_set_affine_lp8(2))
    yield
    A.N.check(lib.sf_set_affine_lp8(0))


@pytest.mark.parametrize("mode", MODES)
def test_affine_scores_4_lane_kernel(mode, lanes4):
    # Band widths over every 4-lane shape (36 .. 72 slots a lane, up to 288 slots), random and
    # gap-heavy pairs, lengths around the value limit (m near 5,200 for ONT at band 16).
    rng = random.Random(61 + int(mode))
    P, T, B = [], [], []
    nmax = 140 if mode == A.Mode.Infix else 900
    for _ in range(500):
        n = rng.randrange(0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.1 else rng.randrange(4) for x in t]
        for _ in range(rng.randrange(0, 4)):
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(0, 9))] + p[k + rng.randrange(0, 9):]
        P.append(p)
        T.append(t)
        # (band 128-131: 66 slots a lane; 132-135: 68; 136-139: 70; 140-143: 72, for m close to n)
        B.append(rng.choice([0, 1, 17, 40, 64, 71, 72, 80, 100, 111, 112, 127, 128, 130, 131, 132, 135, 136,
                             138, 139, 140, 143]))
    if mode != A.Mode.Infix:
        for m in (5150, 5250, 5350):
            t = [rng.randrange(4) for _ in range(m)]
            P.append([x if rng.random() > 0.05 else rng.randrange(4) for x in t])
            T.append(t)
            B.append(16)
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.ONT, A.Affine.HIFI, A.Affine(0, 2, 0, 1), A.Affine(3, 2, 7, 1)):
        check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=False)


@pytest.mark.parametrize("mode", MODES)
def test_affine_traceback_32_lane_kernel(mode, lanes8):
    # CIGAR runs of bands over 512 slots in their fewest-lanes shape: the 32-lane packed traceback
    # kernel (two pairs a wave, 18 to 32 slots a lane), next to the oracle's alignment.
    rng = random.Random(71 + int(mode))
    P, T, B = [], [], []
    nmax = 420 if mode == A.Mode.Infix else 1400  # Infix is unbanded: m + n + 1 slots a row
    for _ in range(160):
        n = rng.randrange(260 if mode == A.Mode.Infix else 0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.1 else rng.randrange(4) for x in t]
        if rng.random() < 0.3:
            k = rng.randrange(len(p) + 1)
            p = p[:k] + [rng.randrange(4) for _ in range(rng.randrange(1, 30))] + p[k + rng.randrange(0, 30):]
        P.append(p)
        T.append(t)
        B.append(rng.choice([256, 257, 300, 400, 500, 511]))
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.ONT, A.Affine.HIFI):
        check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=True)


@pytest.mark.parametrize("mode", MODES)
def test_affine_scores_16bit_gap_corners(mode):
    # The 16-bit kernel lets a vertical gap open only from H before the row's horizontal gaps (a
    # horizontal gap never turns straight into a vertical one), which leaves every H unchanged when
    # mis = 2 ext: the turn ties with cutting its corner.  Pairs built from adjacent deletion and
    # insertion runs, under scorings where those ties are exact (open = 0, mat = 0), against the
    # oracle.
    rng = random.Random(61 + int(mode))
    P, T, B = [], [], []
    nmax = 300 if mode == A.Mode.Infix else 700
    for _ in range(400):
        n = rng.randrange(0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = list(t)
        for _ in range(rng.randrange(0, 6)):
            k = rng.randrange(len(p) + 1)
            a, b = rng.randrange(1, 9), rng.randrange(1, 9)
            p = p[:k] + [rng.randrange(4) for _ in range(b)] + p[k + a:]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 2, 5, 16, 40, 120]))
    ps = A.PairSet.from_seqs(P, T)
    for sc in (A.Affine.ONT, A.Affine(0, 2, 0, 1), A.Affine(1, 4, 0, 2), A.Affine(3, 2, 7, 1)):
        check_affine(ps, mode, sc, np.array(B, np.uint64), cigar=False)


def test_affine_edges():
    P, T = edge_pairs()
    ps = A.PairSet.from_seqs(P[:60], T[:60])
    for mode in MODES:
        for band in (0, 3, 64, None):
            check_affine(ps, mode, A.Affine.ONT, band)


def test_affine_reference_examples():
    # align.rs:3185-3196 through the device
    p = [0, 1, 2, 3, 0, 1, 2, 3]
    assert A.affine(p, p, A.Mode.Global) == 16
    t = [0, 1, 2, 3, 0, 1, 2, 3, 2, 2]
    assert A.affine(p, t, A.Mode.Global) == 16 - 8
    assert A.affine(p, t, A.Mode.Prefix) == 16
    assert A.affine([], [0, 1], A.Mode.Global) == -8
    assert A.affine([0, 1], [], A.Mode.Global) == -8
    assert A.affine_sorted(p, [p, t], A.Mode.Global) == [16, 8]


def test_batch_object_reruns_are_identical():
    ps, band = W.generate(1, 512, 3)
    b = A.Batch(ps)
    b.myers(A.Mode.Infix, locate=2)
    d1 = b.myers_results(2)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=band, cigar=True)
    r1 = b.affine_results(cigar=True)
    b.myers(A.Mode.Infix, locate=2)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=band, cigar=True)
    d2 = b.myers_results(2)
    r2 = b.affine_results(cigar=True)
    for x, y in zip(d1, d2):
        assert np.array_equal(x, y)
    for key in r1:
        assert np.array_equal(r1[key], r2[key])
    t = b.timing()
    assert t.myers_ms > 0 and t.affine_ms > 0 and t.affine_cells > 0
    b.close()


def test_affine_cell_count_is_dp_work():
    # The GCUPS numerator: cells per affine run equal the reference's DP_WORK count (align.rs:
    # 2573-2578) as the oracle restates it, over modes, bands and shapes.
    rng = random.Random(61)
    P, T, B = [], [], []
    for _ in range(200):
        n = rng.randrange(0, 900)
        T.append([rng.randrange(4) for _ in range(n)])
        P.append([rng.randrange(4) for _ in range(max(0, n + rng.randrange(-200, 200)))])
        B.append(rng.choice([0, 3, 64, 128, 700, A.USIZE_MAX]))
    ps = A.PairSet.from_seqs(P, T)
    bands = np.array(B, np.uint64)
    for mode in MODES:
        b = A.Batch(ps)
        b.affine(mode, A.Affine.ONT, bands=bands)
        b.sync()
        want = sum(O.affine_cells(len(P[i]), len(T[i]), int(mode), None if B[i] == A.USIZE_MAX else B[i])
                   for i in range(len(P)))
        assert b.timing().affine_cells == want, mode
        b.close()


def test_myers_long_patterns_class_boundaries():
    # Patterns past the lane kernel (> 24 blocks, 1536 bases) run 16 lanes per pair up to 96
    # blocks and 64 beyond; lengths at the edges of the block classes, texts short enough for the
    # oracle.
    rng = random.Random(41)
    P, T = [], []
    for m in (1536, 1537, 2048, 2049, 3072, 3073, 6144, 6145, 8192, 8193, 12289, 16 * 64 * 24,
              16 * 64 * 24 + 1, 30000, 64 * 64 * 24):
        n = rng.choice([0, 1, 150, 300])
        t = [rng.randrange(4) for _ in range(n)]
        p = [rng.randrange(4) for _ in range(m)]
        if n and rng.random() < 0.5:
            p[:n] = t
        P.append(p)
        T.append(t)
    ps = A.PairSet.from_seqs(P, T)
    for mode in MODES:
        for k in (U32, 2100):
            check_myers(ps, mode, k)


@pytest.mark.parametrize("mode", MODES)
def test_myers_32_lane_pipeline(mode):
    # A bin of 129-160 (or 193-224) block patterns with pairs enough to fill the chip (2,048 at 32
    # lanes a pair) runs the 32-lane pipeline, two pairs a wave, 5 (or 7) blocks a lane; smaller
    # bins run 64 lanes.  Short texts keep the oracle quick.
    rng = np.random.default_rng(83 + int(mode))
    lo, hi = (8193, 10240) if mode != A.Mode.Prefix else (12289, 14336)
    P, T = [], []
    for i in range(2100):
        m = int(rng.integers(lo, hi + 1))
        n = int(rng.integers(0, 90))
        t = rng.integers(0, 4, n, dtype=np.uint8)
        p = rng.integers(0, 4, m, dtype=np.uint8)
        if n and i % 2:
            p[m - n:] = t  # the text matches the pattern's end: small distances too
        P.append(p)
        T.append(t)
    ps = A.PairSet.from_seqs(P, T)
    check_myers(ps, mode, U32)


@pytest.mark.parametrize("mode", MODES)
def test_myers_long_workload(mode):
    ps, _ = W.generate(4, 24, 5)
    check_myers(ps, mode, U32)


def test_issue_probe_runs_every_kind():
    # The VALU issue-rate probe behind the rooflines: every kind runs and takes measurable time;
    # an unknown kind is refused.
    import ctypes as C
    lib = A.N.load_diag()
    ms = C.c_float()
    for kind in range(9):
        A.N.check_diag(lib.sf_probe_issue(kind, 200, 64, C.byref(ms)))
        assert ms.value > 0
    assert lib.sf_probe_issue(9, 200, 64, C.byref(ms)) != 0
    assert lib.sf_probe_issue(-1, 200, 64, C.byref(ms)) != 0


# Scoring values past the 16-bit range (the reference takes any i32 values, `align.rs:660-691`):
# the 32-bit kernels take them up to 2^22 - 1 while the pair's scores stay in range.
LARGE = [A.Affine(300_000, 700_000, 1_000_000, 200_000), A.Affine(4_194_303, 1, 3, 2), A.Affine(7, 4_000_000, 0, 1)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sc", LARGE, ids=["mixed", "mat-max", "mis-max"])
def test_affine_large_scoring_values(mode, sc):
    rng = random.Random(77 + int(mode))
    # Lengths within the library's range bound, (mat + mis + ext)(m + n + 2) < 2^28.
    nmax = min(45, ((1 << 28) // (sc.mat + sc.mis + sc.ext) - 2) // 2 - 6)
    P, T, B = [], [], []
    for _ in range(120):
        n = rng.randrange(0, nmax)
        t = [rng.randrange(4) for _ in range(n)]
        p = [x if rng.random() > 0.1 else rng.randrange(4) for x in t][: rng.randrange(0, n + 1) if n else 0]
        if rng.random() < 0.3:
            p = p + [rng.randrange(4) for _ in range(rng.randrange(0, 6))]
        P.append(p)
        T.append(t)
        B.append(rng.choice([0, 2, 8, A.USIZE_MAX]))
    ps = A.PairSet.from_seqs(P, T)
    check_affine(ps, mode, sc, np.array(B, np.uint64))
    # The passes with bookkeeping (X-drop, best cell, cut) with the same values.
    for i in range(0, len(P), 7):
        sct = (sc.mat, sc.mis, sc.open, sc.ext)
        for clip, cut in (((1_500_000, True), None), ((0, False), 5), (None, 3)):
            want = O.affine_dp(P[i], T[i], int(mode), sct, None if B[i] == A.USIZE_MAX else int(B[i]), None, clip, cut)
            got = A.affine_dp(P[i], T[i], mode, sc, int(B[i]), clip, cut=A.USIZE_MAX if cut is None else cut)
            assert tuple(got) == (want.score, want.score_col, want.clip_best, want.clip_i, want.clip_j,
                                  want.text_read, want.last_row, want.cut_best, want.cut_score), (i, clip, cut)


def test_scoring_values_past_the_bound_are_refused():
    ps = A.PairSet.from_seqs([[0, 1]], [[0, 1]])
    with pytest.raises(A.N.NativeError) as e:
        A.affine_batch(ps, A.Mode.Global, A.Affine(1 << 22, 1, 1, 1), 4)
    assert e.value.status == A.N.SF_ERANGE
