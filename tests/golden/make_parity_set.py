"""Writes tests/golden/parity_set_4096_seed0.npz: the parity-set workload (4,096 pairs from seed 0:
targets uniform ACGT of 200-1,000 bp, queries at 8% error, band 64) with the oracle's results.

    python tests/golden/make_parity_set.py

Per pair: CRC-32s of the pattern and text codes (they pin the generator), the lengths, the band,
Myers Infix (distance, end, start; oracle.locate), affine Global ONT at the pair's band (score,
score column, text start; oracle.align) and the CIGAR (BAM u32 ops, concatenated, with offsets).
The oracle is the C restatement of the reference's kernels (oracle/sf_oracle.c), itself pinned by
the reference's known answers and naive definitions in tests/test_oracle_golden.py.
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import sf_oracle as O  # noqa: E402
from spearfish_amd import workload as W  # noqa: E402

OUT = os.path.join(HERE, "parity_set_4096_seed0.npz")
N_PAIRS, SEED = 4096, 0


def crc(a) -> int:
    return zlib.crc32(np.ascontiguousarray(a, dtype=np.uint8).tobytes())


def main():
    ps, bands = W.generate(1, N_PAIRS, SEED)
    n = len(ps)
    cols = {k: np.zeros(n, np.uint32) for k in ("pcrc", "tcrc", "plen", "tlen", "dist", "end", "start", "score_col",
                                                 "tstart")}
    score = np.zeros(n, np.int32)
    cig, off = [], [0]
    for i in range(n):
        p, t = ps.pattern(i), ps.text(i)
        cols["pcrc"][i], cols["tcrc"][i] = crc(p), crc(t)
        cols["plen"][i], cols["tlen"][i] = len(p), len(t)
        cols["dist"][i], cols["end"][i], cols["start"][i] = O.locate(p, t, O.INFIX)
        s, col, ts, c = O.align(p, t, O.GLOBAL, O.ONT, int(bands[i]))
        score[i], cols["score_col"][i], cols["tstart"][i] = s, col, ts
        cig.extend(c)
        off.append(len(cig))
    np.savez_compressed(OUT, band=bands.astype(np.uint32), score=score, cigar=np.asarray(cig, np.uint32),
                        cigar_off=np.asarray(off, np.uint32), **cols)
    print(f"{OUT}: {n} pairs, {len(cig)} CIGAR ops, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
