"""The headline workload at full size: the 27-point stencil on a 128^3 grid,
rank 0 of 4, C = A_ * P_ in FP64 (524,288 x 557,056, 63.1 M entries), the
product bench.py times.  Checks that do not need the oracle's full product:

* row pointer and every column bit-exact against the closed form of the
  stencil square: row i = (x, y, z) holds the columns i + dz m^2 + dy m + dx
  for every offset in [-2, 2]^3 that stays on the grid, ascending;
* values of 2000 seeded rows against the oracle (componentwise 1e-12);
* C x = A_ (P_ x) for a random x over all rows (SpMV on the host), relative
  to |A_| (|P_| |x|), 1e-13.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def stencil_square_pattern(m, rows, r0=0):
    """Row pointer and global columns of the square of the 27-point stencil
    for rows [r0, r0 + rows) of an m^3 grid (offsets in lexicographic (dz, dy,
    dx) order are ascending in column)."""
    i = np.arange(r0, r0 + rows, dtype=np.int64)
    x, y, z = i % m, (i // m) % m, i // (m * m)
    offs = [(dz, dy, dx) for dz in range(-2, 3) for dy in range(-2, 3) for dx in range(-2, 3)]
    cols = np.empty((rows, len(offs)), np.int32)
    ok = np.empty((rows, len(offs)), bool)
    for q, (dz, dy, dx) in enumerate(offs):
        cols[:, q] = (i + dz * m * m + dy * m + dx).astype(np.int32)
        ok[:, q] = (x + dx >= 0) & (x + dx < m) & (y + dy >= 0) & (y + dy < m) & (z + dz >= 0) & (z + dz < m)
    rpt = np.zeros(rows + 1, np.int64)
    np.cumsum(ok.sum(axis=1), out=rpt[1:])
    return rpt, cols[ok]


def test_stencil128_rank0_share_full_size(orc):
    import scipy.sparse as sp
    import torch
    import nsp_amd
    import bench
    print("case: stencil 128^3 rank-0 share, full size", flush=True)
    A, P, _ = bench.build_c2(128, seed=0)
    dA, dP = nsp_amd.DeviceCSR.from_host(A), nsp_amd.DeviceCSR.from_host(P)
    C = nsp_amd.Handle().spgemm(dA, dP)
    assert (C.nrows, C.ncols) == (A.nrows, P.ncols)
    rpt_ref, col_ref = stencil_square_pattern(128, A.nrows)
    assert C.nnz == col_ref.size == 63_107_092
    np.testing.assert_array_equal(C.row_ptr.cpu().numpy().astype(np.int64), rpt_ref)
    col = C.col.cpu().numpy()
    assert np.array_equal(col, col_ref), "columns differ from the closed form"
    del col_ref
    val = C.val.cpu().numpy()
    # 2000 seeded rows against the oracle
    rows = np.unique(np.random.default_rng(11).choice(A.nrows, 2000, replace=False)).astype(np.int64)
    r_rpt, r_col, r_val, r_abs = orc.spgemm_rows(A, P, rows, with_abs=True)
    got = np.concatenate([val[rpt_ref[i]:rpt_ref[i + 1]] for i in rows])
    got_col = np.concatenate([col[rpt_ref[i]:rpt_ref[i + 1]] for i in rows])
    np.testing.assert_array_equal(got_col, r_col)
    comp, _ = orc.componentwise_error(got, r_val, r_abs)
    assert comp <= 1e-12, comp
    # C x = A_ (P_ x) over every row
    Cs = sp.csr_matrix((val, col, rpt_ref), shape=(C.nrows, C.ncols))
    As = sp.csr_matrix((A.val, A.col, A.rpt), shape=(A.nrows, A.ncols))
    Ps = sp.csr_matrix((P.val, P.col, P.rpt), shape=(P.nrows, P.ncols))
    x = np.random.default_rng(12).uniform(-1, 1, P.ncols)
    y1, y2 = Cs @ x, As @ (Ps @ x)
    scale = abs(As) @ (abs(Ps) @ np.abs(x))
    rel = float(np.max(np.abs(y1 - y2) / np.where(scale > 0, scale, 1.0)))
    print(f"  nnz {C.nnz}, sampled rows err {comp:.2e}, C x vs A(P x) err {rel:.2e}", flush=True)
    assert rel <= 1e-13, rel
    del C
    torch.cuda.empty_cache()


@pytest.mark.parametrize("rank", [0, 1])
def test_stencil128_share_compaction_full_size(orc, rank):
    """The same 128^3 stencil split 4 ways, rank 0 and rank 1, through the
    whole per-rank path at full size (spspmpi.cu:371-975 without MPI): A's
    row block with rank-frame columns, halo discovery and A's compaction
    (nsp_shrink_cols), the completedP gather from the full P
    (nsp_gather_rows_*), P's compaction, then the product.  Rank 1's block
    has halo planes on both sides, so neither map is the identity.  A_.col,
    completedP (row pointer, columns, values) and P's compaction map are
    compared bit-exactly with the oracle's local share; C's columns, mapped
    back through P's map to global columns, equal the closed form of the
    stencil square on rows [r0, r1); 2000 seeded rows of values against the
    oracle; C x = A_ (P_ x) over every row."""
    import scipy.sparse as sp
    import torch
    import nsp_amd
    from nsp_amd import csr, gen
    print("case: stencil 128^3 share compaction, rank", rank, flush=True)
    m, nprocs = 128, 4
    n = m ** 3
    r0, r1 = csr.rank_rows(n, nprocs, rank)
    P = gen.stencil_27pt(m, seed=0)
    A_rank = P.row_slice(r0, r1, col_shift=r0)
    ref = orc.local_share(A_rank, P, r0)
    Al = csr.CSR.from_device(nsp_amd.DeviceCSR.from_host(A_rank), full_n=n, row_shift=r0)
    Al.m = n
    dP = nsp_amd.DeviceCSR.from_host(P)
    C, A_, P_, loc_time = csr.nsparse_mgpu_version1_local(Al, dP)
    np.testing.assert_array_equal(A_.col.cpu().numpy(), ref["A_"].col)
    assert A_.m == ref["A_"].ncols
    np.testing.assert_array_equal(P_.row.cpu().numpy().astype(np.int64), ref["P_"].rpt)
    np.testing.assert_array_equal(P_.col.cpu().numpy(), ref["P_"].col)
    np.testing.assert_array_equal(P_.val.cpu().numpy(), ref["P_"].val)
    perm_p = C.extra["shrinking_permut_P"]
    np.testing.assert_array_equal(perm_p.cpu().numpy(), ref["perm_P"])
    assert P_.m == ref["P_"].ncols
    del dP
    print(f"  A_ {A_.n} x {A_.m}, completedP {P_.n} x {P_.m}, maps and completedP bit-exact", flush=True)
    # structure: compacted columns -> global columns, against the closed form
    rpt_ref, col_ref = stencil_square_pattern(m, r1 - r0, r0)
    assert C.nnz == col_ref.size
    np.testing.assert_array_equal(C.row.cpu().numpy().astype(np.int64), rpt_ref)
    gcol = (perm_p.to(torch.int64)[C.col.to(torch.int64)] + r0).cpu().numpy()
    assert np.array_equal(gcol, col_ref.astype(np.int64)), "global columns differ from the closed form"
    del gcol, col_ref
    col = C.col.cpu().numpy()
    val = C.val.cpu().numpy()
    rows = np.unique(np.random.default_rng(13 + rank).choice(r1 - r0, 2000, replace=False)).astype(np.int64)
    r_rpt, r_col, r_val, r_abs = orc.spgemm_rows(ref["A_"], ref["P_"], rows, with_abs=True)
    got = np.concatenate([val[rpt_ref[i]:rpt_ref[i + 1]] for i in rows])
    got_col = np.concatenate([col[rpt_ref[i]:rpt_ref[i + 1]] for i in rows])
    np.testing.assert_array_equal(got_col, r_col)
    comp, _ = orc.componentwise_error(got, r_val, r_abs)
    assert comp <= 1e-12, comp
    Cs = sp.csr_matrix((val, col, rpt_ref), shape=(C.n, C.m))
    As = sp.csr_matrix((ref["A_"].val, ref["A_"].col, ref["A_"].rpt), shape=(ref["A_"].nrows, ref["A_"].ncols))
    Ps = sp.csr_matrix((ref["P_"].val, ref["P_"].col, ref["P_"].rpt), shape=(ref["P_"].nrows, ref["P_"].ncols))
    x = np.random.default_rng(14).uniform(-1, 1, C.m)
    y1, y2 = Cs @ x, As @ (Ps @ x)
    scale = abs(As) @ (abs(Ps) @ np.abs(x))
    rel = float(np.max(np.abs(y1 - y2) / np.where(scale > 0, scale, 1.0)))
    print(f"  nnz {C.nnz}, sampled rows err {comp:.2e}, C x vs A_(P_ x) err {rel:.2e}, loc_time {loc_time:.4f} s",
          flush=True)
    assert rel <= 1e-13, rel
    del C, A_, P_
    torch.cuda.empty_cache()
