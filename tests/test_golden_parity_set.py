"""The committed parity-set fixture (tests/golden/parity_set_4096_seed0.npz, written by
tests/golden/make_parity_set.py): 4,096 pairs from seed 0, Myers Infix (distance, end, start) and
affine Global ONT at band 64 with CIGAR.

CPU: the generator still makes the same pairs (CRC-32 per sequence) and the oracle still gives the
fixture's results.  GPU: every pair through the C ABI equals the fixture, bit for bit.
"""
import os
import zlib

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import align as A
from spearfish_amd import workload as W

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "parity_set_4096_seed0.npz")


@pytest.fixture(scope="module")
def golden():
    g = dict(np.load(FIXTURE))
    ps, bands = W.generate(1, 4096, 0)
    return g, ps, bands


def cigar_of(g, i):
    return [int(x) for x in g["cigar"][int(g["cigar_off"][i]):int(g["cigar_off"][i + 1])]]


def test_generator_makes_the_fixture_pairs(golden):
    g, ps, bands = golden
    assert len(ps) == 4096
    assert np.array_equal(bands.astype(np.uint32), g["band"])
    for i in range(len(ps)):
        p, t = ps.pattern(i), ps.text(i)
        assert (len(p), len(t)) == (g["plen"][i], g["tlen"][i]), i
        assert zlib.crc32(p.tobytes()) == g["pcrc"][i] and zlib.crc32(t.tobytes()) == g["tcrc"][i], i


def test_oracle_reproduces_the_fixture(golden):
    g, ps, bands = golden
    for i in range(0, len(ps), 3):
        p, t = ps.pattern(i), ps.text(i)
        assert O.locate(p, t, O.INFIX) == (g["dist"][i], g["end"][i], g["start"][i]), i
        s, col, ts, c = O.align(p, t, O.GLOBAL, O.ONT, int(bands[i]))
        assert (s, col, ts) == (g["score"][i], g["score_col"][i], g["tstart"][i]), i
        assert c == cigar_of(g, i), i


@pytest.mark.gpu
def test_device_equals_the_fixture_on_every_pair(golden):
    g, ps, bands = golden
    b = A.Batch(ps)
    b.myers(A.Mode.Infix, locate=2)
    dist, end, start = b.myers_results(2)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=True)
    r = b.affine_results(cigar=True)
    b.close()
    assert np.array_equal(dist, g["dist"]) and np.array_equal(end, g["end"]) and np.array_equal(start, g["start"])
    assert np.array_equal(r["score"], g["score"]) and np.array_equal(r["score_col"], g["score_col"])
    assert np.array_equal(r["tstart"], g["tstart"])
    assert np.array_equal(r["cigar_off"].astype(np.uint32), g["cigar_off"])
    assert np.array_equal(r["cigar"], g["cigar"])
