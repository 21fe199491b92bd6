"""sf_myers_noalloc against sf_myers through the C ABI: the same distances, ends and start columns,
bit for bit, on two calls in a row (the second smaller, so stale data from the first in the shared
device array would show), in all three modes, packed and unpacked input, and every Myers class up
to the 64-lane pipelines; inputs over its 32 MiB limit are refused with SF_ERANGE."""
import ctypes as C

import numpy as np
import pytest

from spearfish_amd import _native as N
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu


def run(fn, ps, mode, k):
    n = len(ps)
    d, e, s = (np.full(n, 7, np.uint32) for _ in range(3))
    cp = ps.c()
    st = getattr(N.load(), fn)(C.byref(cp), int(mode), k, d.ctypes.data, e.ctypes.data, s.ctypes.data)
    return st, d, e, s


def mixed_pairs(seed, n):
    """Patterns of every Myers class (1 to 1,536 blocks over the set), texts 0.9-1.3x longer."""
    rng = np.random.default_rng(seed)
    lens = np.concatenate([rng.integers(0, 1600, n - 12), [0, 1, 64, 65, 1536, 1537, 6144, 6145, 24576, 24577,
                                                            30000, 40000]])
    P, T = [], []
    for m in lens:
        t = rng.integers(0, 4, int(m * rng.uniform(0.9, 1.3)) + int(rng.integers(0, 40)), dtype=np.uint8)
        p = t[:m].copy() if len(t) >= m else rng.integers(0, 4, m, dtype=np.uint8)
        flip = rng.random(len(p)) < 0.08
        p[flip] = (p[flip] + 1) % 4
        P.append(p)
        T.append(t)
    return A.PairSet.from_seqs(P, T)


def packed(ps):
    def pack(codes):
        w = np.zeros((len(codes) + 15) // 16 + 1, np.uint32)
        k = np.arange(len(codes))
        np.bitwise_or.at(w, k >> 4, (codes.astype(np.uint32) & 3) << (2 * (k & 15)).astype(np.uint32))
        return w
    return A.PairSet(pack(ps.pat), ps.pat_off, ps.pat_len, pack(ps.txt), ps.txt_off, ps.txt_len, packed=True)


@pytest.mark.parametrize("mode", [A.Mode.Infix, A.Mode.Prefix, A.Mode.Global])
@pytest.mark.parametrize("k", [A.U32_MAX, 50], ids=["k=inf", "k=50"])
def test_noalloc_equals_sf_myers_twice_second_smaller(mode, k):
    first, _ = W.generate(1, 3000, 11)
    second = mixed_pairs(int(mode) * 7 + (k == 50), 300)
    for ps in (first, second, packed(second)):
        st0, d0, e0, s0 = run("sf_myers", ps, mode, k)
        st1, d1, e1, s1 = run("sf_myers_noalloc", ps, mode, k)
        assert st0 == N.SF_OK and st1 == N.SF_OK, N.load().sf_last_error()
        assert np.array_equal(d0, d1) and np.array_equal(e0, e1) and np.array_equal(s0, s1)


def test_noalloc_limit_and_null_outputs():
    # Unpacked codes take 1.25 B a base in the array: 30,000 pairs of 1,000 + ~1,000 bases need
    # about 75 MB and are refused; 10,000 need about 26 MB and run (end and start left null).
    big, _ = W.generate(2, 30000, 3)
    st, *_ = run("sf_myers_noalloc", big, A.Mode.Infix, A.U32_MAX)
    assert st == N.SF_ERANGE and b"32 MiB" in N.load().sf_last_error()
    ps, _ = W.generate(2, 10000, 3)
    d0 = np.empty(len(ps), np.uint32)
    d1 = np.empty(len(ps), np.uint32)
    lib = N.load()
    cp = ps.c()
    assert lib.sf_myers(C.byref(cp), 0, A.U32_MAX, d0.ctypes.data, None, None) == N.SF_OK
    assert lib.sf_myers_noalloc(C.byref(cp), 0, A.U32_MAX, d1.ctypes.data, None, None) == N.SF_OK
    assert np.array_equal(d0, d1)
