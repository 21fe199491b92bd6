"""The autotuner: measuring the uneven workload's candidate shape classes on the device changes only
how pairs are laid out on it, never a result; the measurements round-trip through a file."""
import os

import numpy as np
import pytest

from oracle import sf_oracle as O
from spearfish_amd import _native as N
from spearfish_amd import align as A
from spearfish_amd import workload as W

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results_and_persists(tmp_path):
    lib = N.load()
    N.check(lib.sf_tune_reset())
    ps, bands = W.generate(3, 30000, 6)
    b = A.Batch(ps)
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    before = b.affine_results()
    b.autotune(A.Mode.Global, A.Affine.ONT, bands=bands)
    n = lib.sf_tune_classes()
    assert n >= 4
    b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    after = b.affine_results()
    assert np.array_equal(before["score"], after["score"]) and np.array_equal(before["score_col"], after["score_col"])
    for i in range(0, len(ps), 997):
        dp = O.affine_dp(ps.pattern(i), ps.text(i), O.GLOBAL, O.ONT, int(bands[i]))
        assert (int(after["score"][i]), int(after["score_col"][i])) == (dp.score, dp.score_col), i
    path = str(tmp_path / "tune.txt")
    N.check(lib.sf_tune_save(path.encode()))
    lines = [x for x in open(path).read().split("\n") if x]
    assert all(x.startswith("gfx950-") for x in lines)
    assert len([x for x in lines if " setting " not in x]) == n
    assert len([x for x in lines if " setting " in x]) == 1  # the planner setting the search kept
    N.check(lib.sf_tune_reset())
    assert lib.sf_tune_classes() == 0
    N.check(lib.sf_tune_load(path.encode()))
    assert lib.sf_tune_classes() == n
    assert lib.sf_tune_load(str(tmp_path / "missing.txt").encode()) == N.SF_EINVAL
    N.check(lib.sf_tune_reset())
    b.close()
