"""Property tests (hypothesis) of the oracle's product machinery on random
small integer matrices, against dense brute force: empty matrices, empty
rows and columns, one-entry rows, rows selecting many B rows.

  * C = A*B: structure = nonzeros of pattern(A) @ pattern(B) (zeros kept
    where products cancel), rows ascending, values = the exact integer
    product; symbolic row counts = the structure's row lengths;
  * intprod_i = sum over A_i's entries of nnz(B_k), total = their sum;
  * transpose = the dense transpose, and transposing twice is the identity;
  * exscan = numpy cumsum with a leading 0;
  * shrink(col, l): perm = sorted unique(col U [0..l]), perm[newcol] = col.
"""
import numpy as np
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from nsp_amd.gen import HostCSR

SETTINGS = settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])


@st.composite
def dense_int(draw, rows=None, cols=None, max_dim=20):
    r = draw(st.integers(0, max_dim)) if rows is None else rows
    c = draw(st.integers(0, max_dim)) if cols is None else cols
    density = draw(st.floats(0.0, 0.6))
    seed = draw(st.integers(0, 2**31 - 1))
    rng = np.random.default_rng(seed)
    D = np.where(rng.random((r, c)) < density, rng.integers(1, 4, (r, c)) * rng.choice([-1, 1], (r, c)), 0)
    return D.astype(np.int64)


def csr(D):
    rows, cols = np.nonzero(D)
    rpt = np.zeros(D.shape[0] + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=D.shape[0]), out=rpt[1:])
    return HostCSR(rpt, cols.astype(np.int32), D[rows, cols].astype(np.float64), D.shape[0], D.shape[1])


@st.composite
def pair(draw):
    m, k, n = (draw(st.integers(0, 20)) for _ in range(3))
    return draw(dense_int(m, k)), draw(dense_int(k, n))


@SETTINGS
@given(pair())
def test_product_is_dense_product(orc, ab):
    Ad, Bd = ab
    A, B = csr(Ad), csr(Bd)
    rpt, col, val, cabs = orc.spgemm(A, B)
    pat = ((Ad != 0).astype(np.int64) @ (Bd != 0).astype(np.int64)) > 0
    exact = Ad @ Bd
    rows, cols = np.nonzero(pat)
    np.testing.assert_array_equal(np.diff(rpt), pat.sum(axis=1))
    np.testing.assert_array_equal(col, cols)
    np.testing.assert_array_equal(val, exact[rows, cols].astype(np.float64))
    np.testing.assert_array_equal(cabs, (np.abs(Ad) @ np.abs(Bd))[rows, cols].astype(np.float64))
    np.testing.assert_array_equal(orc.symbolic(A, B), pat.sum(axis=1))


@SETTINGS
@given(pair())
def test_intprod_counts_b_rows(orc, ab):
    Ad, Bd = ab
    A, B = csr(Ad), csr(Bd)
    ip, total = orc.intprod(A, B)
    expect = (Ad != 0).astype(np.int64) @ (Bd != 0).sum(axis=1)
    np.testing.assert_array_equal(ip, expect)
    assert total == int(expect.sum())


@SETTINGS
@given(dense_int())
def test_transpose_is_dense_transpose(orc, D):
    X = csr(D)
    trpt, tcol, tval = orc.transpose(X)
    T = csr(D.T.copy())
    np.testing.assert_array_equal(trpt, T.rpt)
    np.testing.assert_array_equal(tcol, T.col)
    np.testing.assert_array_equal(tval, T.val)
    rrpt, rcol, rval = orc.transpose(HostCSR(trpt, tcol, tval, D.shape[1], D.shape[0]))
    np.testing.assert_array_equal(rrpt, X.rpt)
    np.testing.assert_array_equal(rcol, X.col)
    np.testing.assert_array_equal(rval, X.val)


@SETTINGS
@given(st.lists(st.integers(0, 10**6), max_size=300))
def test_exscan_is_cumsum(orc, xs):
    x = np.array(xs, np.int64)
    out = orc.exscan(x)
    np.testing.assert_array_equal(out, np.concatenate([[0], np.cumsum(x)]).astype(np.int64))


@SETTINGS
@given(st.lists(st.integers(-50, 200), max_size=200), st.integers(0, 60))
def test_shrink_is_unique_and_search(orc, cols, l):
    col = np.array(cols, np.int32)
    perm, newcol = orc.shrink(col, l)
    expect = np.unique(np.concatenate([col, np.arange(l + 1, dtype=np.int32)]))
    np.testing.assert_array_equal(perm, expect)
    np.testing.assert_array_equal(perm[newcol], col)
