"""The oracle against the committed fixtures of the shell scene's 64x64 view
(tests/golden/shell64.npz, written by tests/golden/make_fixtures.py): hit
flags, sorted hit rows and coarse sample streams bit for bit, the FP64 image
to rounding.  A regression anchor: a change of the oracle's arithmetic or
ordering shows up here."""
import os

import numpy as np

import make_fixtures as F

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_reproduces_the_fixtures():
    want = np.load(os.path.join(HERE, "golden", "shell64.npz"))
    got = F.fixtures()
    for k in ("hit", "rays", "idx", "lo", "hi", "c_idx", "c_depth", "c_dist"):
        assert np.array_equal(got[k], want[k]), k
    for k in ("rgb", "depth", "missed"):
        assert np.abs(got[k] - want[k]).max() < 1e-10, k
