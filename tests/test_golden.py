"""Known-answer fixtures (tests/golden/*.npz, written by make_fixtures.py
from exact dense integer products, structure from the boolean pattern
product, zeros kept where products cancel) against the oracle: row pointer,
columns and values equal."""
import glob
import os

import numpy as np
import pytest

from nsp_amd.gen import HostCSR

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    z = np.load(path)
    m, k, n = (int(x) for x in z["shape"])
    A = HostCSR(z["a_rpt"], z["a_col"].astype(np.int32), z["a_val"], m, k)
    B = HostCSR(z["b_rpt"], z["b_col"].astype(np.int32), z["b_val"], k, n)
    return A, B, (z["c_rpt"], z["c_col"], z["c_val"])


def test_fixtures_present():
    assert len(FIXTURES) >= 6


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_fixture(orc, path):
    A, B, (crpt, ccol, cval) = load(path)
    rpt, col, val, _ = orc.spgemm(A, B)
    np.testing.assert_array_equal(rpt, crpt)
    np.testing.assert_array_equal(col, ccol)
    np.testing.assert_array_equal(val, cval)
