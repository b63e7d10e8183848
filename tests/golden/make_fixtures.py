"""Writes the small known-answer fixtures of the local product C = A*B
(tests/golden/*.npz) from exact dense arithmetic, independently of the
oracle and of the GPU code:

  * values are small integers, so the dense int64 product A @ B is exact and
    equals the FP64 (and FP32) result of any summation order;
  * the structure of C is the union of the B rows each A row selects
    (nsp/src/nsp.cu keeps structural zeros: an entry whose products cancel
    stays in C with value 0), i.e. the nonzeros of the boolean product
    pattern(A) @ pattern(B);
  * rows of C sorted ascending.

Each file holds a_rpt, a_col, a_val, b_rpt, b_col, b_val, c_rpt, c_col,
c_val (int64 / float64 arrays) and shape = (rows A, cols A, cols B).

    python tests/golden/make_fixtures.py        # rewrites the .npz files
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_of_dense(D):
    rpt = np.zeros(D.shape[0] + 1, np.int64)
    rows, cols = np.nonzero(D)
    np.cumsum(np.bincount(rows, minlength=D.shape[0]), out=rpt[1:])
    return rpt, cols.astype(np.int64), D[rows, cols].astype(np.float64)


def product(Ad, Bd):
    """Expected C: structure from the boolean product of the patterns,
    values from the exact integer product (zeros kept where they cancel)."""
    pat = ((Ad != 0).astype(np.int64) @ (Bd != 0).astype(np.int64)) > 0
    vals = Ad.astype(np.int64) @ Bd.astype(np.int64)
    rows, cols = np.nonzero(pat)
    rpt = np.zeros(Ad.shape[0] + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=Ad.shape[0]), out=rpt[1:])
    return rpt, cols.astype(np.int64), vals[rows, cols].astype(np.float64)


def laplacian_5pt(m):
    n = m * m
    D = np.zeros((n, n), np.int64)
    for i in range(n):
        x, y = i % m, i // m
        D[i, i] = 4
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            if 0 <= x + dx < m and 0 <= y + dy < m:
                D[i, (y + dy) * m + x + dx] = -1
    return D


def random_int(rng, r, c, density, empty_rows=()):
    D = np.where(rng.random((r, c)) < density, rng.integers(1, 4, (r, c)) * rng.choice([-1, 1], (r, c)), 0)
    for i in empty_rows:
        D[i, :] = 0
    return D.astype(np.int64)


def cases():
    rng = np.random.default_rng(2024)
    L = laplacian_5pt(8)
    yield "laplacian_8x8_squared", L, L
    A = random_int(rng, 48, 40, 0.15, empty_rows=(0, 17, 47))
    B = random_int(rng, 40, 56, 0.12, empty_rows=(3, 39))
    yield "random_48x40x56", A, B
    # exact cancellation: a_i0 b_0j + a_i1 b_1j = 0 for the first columns
    A = np.zeros((6, 4), np.int64)
    A[:, 0], A[:, 1] = 1, -1
    A[5, :] = 0
    A[2, 3] = 2
    B = np.zeros((4, 9), np.int64)
    B[0, :5] = 3
    B[1, 2:7] = 3
    B[3, ::2] = 1
    yield "cancellation_zeros_kept", A, B
    A = random_int(rng, 1, 64, 0.6)
    B = random_int(rng, 64, 64, 0.25)
    yield "one_long_row", A, B
    A = random_int(rng, 64, 1, 1.0)
    B = random_int(rng, 1, 64, 0.7)
    yield "outer_product", A, B
    A = np.eye(37, dtype=np.int64) * 2
    B = random_int(rng, 37, 29, 0.2, empty_rows=(5,))
    yield "scaled_identity", A, B


def main():
    for name, Ad, Bd in cases():
        a_rpt, a_col, a_val = csr_of_dense(Ad)
        b_rpt, b_col, b_val = csr_of_dense(Bd)
        c_rpt, c_col, c_val = product(Ad, Bd)
        np.savez(os.path.join(HERE, name + ".npz"), a_rpt=a_rpt, a_col=a_col, a_val=a_val, b_rpt=b_rpt,
                 b_col=b_col, b_val=b_val, c_rpt=c_rpt, c_col=c_col, c_val=c_val,
                 shape=np.array([Ad.shape[0], Ad.shape[1], Bd.shape[1]], np.int64))
        print(f"{name}: A {Ad.shape} nnz {a_col.size}, B {Bd.shape} nnz {b_col.size}, C nnz {c_col.size}, "
              f"explicit zeros {int((c_val == 0).sum())}")


if __name__ == "__main__":
    main()
