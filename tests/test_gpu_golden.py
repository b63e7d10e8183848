"""Known-answer fixtures (tests/golden/*.npz: exact dense integer products,
zeros kept where products cancel) against the HIP product in every index
and value type: row pointer, columns and values equal (small integers, so
any summation order is exact in FP64 and FP32)."""
import os

import numpy as np
import pytest

from tests.test_golden import FIXTURES, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("index,dtype", [("i32", "f64"), ("i64", "f64"), ("i32", "f32")])
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_product_matches_fixture(path, index, dtype):
    import nsp_amd
    print("case: fixture", os.path.basename(path), index, dtype, flush=True)
    A, B, (crpt, ccol, cval) = load(path)
    C = nsp_amd.Handle().spgemm(nsp_amd.DeviceCSR.from_host(A, dtype, index),
                                nsp_amd.DeviceCSR.from_host(B, dtype, index))
    np.testing.assert_array_equal(C.row_ptr.cpu().numpy().astype(np.int64), crpt)
    np.testing.assert_array_equal(C.col.cpu().numpy(), ccol)
    np.testing.assert_array_equal(C.val.cpu().numpy().astype(np.float64), cval)
