"""Seeded synthetic CSR workloads (host side, numpy).

These are input generators, not reference code: the reference ships no
matrices and reads them from binary COO files (src/READ_MATRIX_MPI.cpp:146).
Every generator returns a ``HostCSR`` with sorted, duplicate-free rows,
int64 row pointers, int32 columns and float64 values, built the same way on
every machine for a given seed.

Workloads:
  * ``laplacian_5pt(m)``          2-D 5-point Laplacian on an m x m grid
  * ``stencil_27pt(m, seed, ...)``  3-D 27-point stencil on m^3, values
                                    U(-1,1) off the diagonal, 27 on it; any
                                    contiguous row block can be generated on
                                    its own and equals that block of the full
                                    matrix
  * ``rmat(scale, seed, ...)``      R-MAT power-law graph, duplicates merged,
                                    self-loops kept, values U(-1,1)
  * ``aggregation(n, seed)``        n x ceil(n/8) prolongator, one entry per
                                    row at column floor(i/8), U(0.5,1.5)
  * ``uneven_batch(count, seed)``   independent square matrices of uneven
                                    sizes for the batched product
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostCSR:
    rpt: np.ndarray  # int64 [nrows+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # float64 [nnz]
    nrows: int
    ncols: int

    @property
    def nnz(self) -> int:
        return int(self.rpt[-1])

    def row_slice(self, r0: int, r1: int, col_shift: int = 0) -> "HostCSR":
        """Rows [r0, r1) as a new CSR; columns shifted by -col_shift."""
        b, e = int(self.rpt[r0]), int(self.rpt[r1])
        col = self.col[b:e]
        if col_shift:
            col = (col.astype(np.int64) - col_shift).astype(np.int32)
        return HostCSR(self.rpt[r0:r1 + 1] - b, np.ascontiguousarray(col),
                       np.ascontiguousarray(self.val[b:e]), r1 - r0, self.ncols)

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols))
        rows = np.repeat(np.arange(self.nrows), np.diff(self.rpt))
        d[rows, self.col] = self.val
        return d


def _from_rows_cols(nrows, ncols, rows, cols, vals) -> HostCSR:
    """CSR from (row, col) pairs that are already unique; sorts rows and columns."""
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rpt = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=nrows), out=rpt[1:])
    return HostCSR(rpt, cols.astype(np.int32), vals.astype(np.float64), nrows, ncols)


def laplacian_5pt(m: int) -> HostCSR:
    """5-point Laplacian on an m x m grid: 4 on the diagonal, -1 to each grid
    neighbour; row id = y*m + x."""
    n = m * m
    i = np.arange(n, dtype=np.int64)
    x, y = i % m, i // m
    offs = [(-m, y > 0), (-1, x > 0), (0, np.ones(n, bool)), (1, x < m - 1), (m, y < m - 1)]
    cols = np.stack([np.where(ok, i + d, -1) for d, ok in offs], axis=1)
    vals = np.stack([np.full(n, 4.0 if d == 0 else -1.0) for d, _ in offs], axis=1)
    mask = cols >= 0
    rpt = np.zeros(n + 1, np.int64)
    np.cumsum(mask.sum(1), out=rpt[1:])
    return HostCSR(rpt, cols[mask].astype(np.int32), vals[mask], n, n)


def stencil_27pt_counts(m: int, r0: int, r1: int) -> np.ndarray:
    """Number of entries in each row of the 27-point stencil, rows [r0, r1)."""
    i = np.arange(r0, r1, dtype=np.int64)
    x, y, z = i % m, (i // m) % m, i // (m * m)

    def span(c):
        return 3 - (c == 0).astype(np.int64) - (c == m - 1).astype(np.int64)

    return span(x) * span(y) * span(z)


def stencil_27pt(m: int, seed: int = 0, r0: int = 0, r1: int | None = None) -> HostCSR:
    """Rows [r0, r1) of the 3-D 27-point stencil on an m^3 grid.

    Off-diagonal values are U(-1,1) drawn from PCG64(seed) in CSR order over
    the *full* matrix; the block starts its draws where the full matrix's
    draws for row r0 start, so any block equals the same rows of the full
    matrix.  The diagonal is 27.  Columns are global (not shifted)."""
    n = m ** 3
    if r1 is None:
        r1 = n
    rows = np.arange(r0, r1, dtype=np.int64)
    x, y, z = rows % m, (rows // m) % m, rows // (m * m)
    cols = np.empty((r1 - r0, 27), np.int64)
    t = 0
    for dz in (-1, 0, 1):
        okz = (z + dz >= 0) & (z + dz < m)
        for dy in (-1, 0, 1):
            oky = okz & (y + dy >= 0) & (y + dy < m)
            for dx in (-1, 0, 1):
                ok = oky & (x + dx >= 0) & (x + dx < m)
                cols[:, t] = np.where(ok, rows + dz * m * m + dy * m + dx, -1)
                t += 1
    mask = cols >= 0
    cnt = mask.sum(1)
    rpt = np.zeros(r1 - r0 + 1, np.int64)
    np.cumsum(cnt, out=rpt[1:])
    colv = cols[mask]
    before = int(stencil_27pt_counts(m, 0, r0).sum()) if r0 > 0 else 0
    bg = np.random.PCG64(seed)
    if before:
        bg.advance(before)
    val = np.random.Generator(bg).uniform(-1.0, 1.0, colv.size)
    diag = colv == np.repeat(rows, cnt)
    val[diag] = 27.0
    return HostCSR(rpt, colv.astype(np.int32), val, r1 - r0, n)


def rmat(scale: int, seed: int = 12345, edge_factor: int = 16,
         a: float = 0.57, b: float = 0.19, c: float = 0.19) -> HostCSR:
    """R-MAT graph with 2^scale vertices and edge_factor*2^scale edges drawn
    by recursive quadrant choice (a, b, c, d=1-a-b-c); duplicates merged,
    self-loops kept, values U(-1,1)."""
    n = 1 << scale
    ne = edge_factor * n
    rng = np.random.default_rng(seed)
    r = np.zeros(ne, np.int64)
    cidx = np.zeros(ne, np.int64)
    for lvl in range(scale):
        u = rng.random(ne)
        down = u >= a + b            # quadrant c or d: row bit set
        right = ((u >= a) & (u < a + b)) | (u >= a + b + c)   # b or d: col bit set
        r |= down.astype(np.int64) << lvl
        cidx |= right.astype(np.int64) << lvl
    key = np.unique(r * n + cidx)
    rows, cols = key // n, key % n
    vals = rng.uniform(-1.0, 1.0, key.size)
    rpt = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rpt[1:])
    return HostCSR(rpt, cols.astype(np.int32), vals, n, n)


def aggregation(n: int, seed: int = 0, agg: int = 8) -> HostCSR:
    """Aggregation prolongator: one entry per row at column floor(i/agg)."""
    rng = np.random.default_rng(seed)
    rpt = np.arange(n + 1, dtype=np.int64)
    col = (np.arange(n) // agg).astype(np.int32)
    val = rng.uniform(0.5, 1.5, n)
    return HostCSR(rpt, col, val, n, (n + agg - 1) // agg)


def random_csr(nrows: int, ncols: int, density: float, seed: int = 0,
               integer_values: bool = False, empty_rows: float = 0.0) -> HostCSR:
    """Uniform random pattern; integer_values draws small nonzero integers so
    dense float64 products of small matrices are exact."""
    rng = np.random.default_rng(seed)
    nnz_target = max(0, int(round(nrows * ncols * density)))
    key = np.unique(rng.integers(0, max(nrows * ncols, 1), nnz_target, dtype=np.int64))
    rows, cols = key // max(ncols, 1), key % max(ncols, 1)
    if empty_rows > 0 and nrows > 0:
        drop = rng.random(nrows) < empty_rows
        keep = ~drop[rows]
        rows, cols = rows[keep], cols[keep]
    if integer_values:
        vals = rng.integers(1, 5, rows.size) * rng.choice([-1.0, 1.0], rows.size)
    else:
        vals = rng.uniform(-1.0, 1.0, rows.size)
    return _from_rows_cols(nrows, ncols, rows, cols, vals.astype(np.float64))


def uneven_batch_concat(count: int = 4096, seed: int = 0, nmin: int = 64, nmax: int = 4096,
                        lam: float = 8.0) -> tuple[HostCSR, np.ndarray]:
    """`count` independent square matrices as one concatenated CSR (rows of
    matrix 0, then matrix 1, ...; columns local to each matrix) and the row
    offsets of the matrices (count+1 entries).  n_i ~ U[nmin, nmax]; entries
    per row ~ max(Poisson(lam), 1) (capped at n_i); columns uniform without
    replacement; values U(-1,1).  The concatenated CSR's ncols is the largest
    n_i."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, count).astype(np.int64)
    row_off = np.zeros(count + 1, np.int64)
    np.cumsum(ns, out=row_off[1:])
    M = int(row_off[-1])
    nrow = np.repeat(ns, ns)                          # n of each row's matrix
    k = np.minimum(np.maximum(rng.poisson(lam, M), 1), nrow)
    rpt = np.zeros(M + 1, np.int64)
    np.cumsum(k, out=rpt[1:])
    S = np.int64(nmax + 1)
    rows = np.repeat(np.arange(M, dtype=np.int64), k)
    key = rows * S + (rng.random(rows.size) * np.repeat(nrow, k)).astype(np.int64)
    del rows
    key.sort()
    # a duplicate (row, col) is dropped and its row draws again; the sorted
    # distinct keys and the sorted redraws are merged (two runs: linear)
    while True:
        dup = np.flatnonzero(key[1:] == key[:-1]) + 1
        if dup.size == 0:
            break
        r = key[dup] // S
        redraw = r * S + (rng.random(dup.size) * nrow[r]).astype(np.int64)
        redraw.sort()
        key = np.sort(np.concatenate([np.delete(key, dup), redraw]), kind="stable")
    cols = (key % S).astype(np.int32)
    del key
    vals = rng.uniform(-1.0, 1.0, cols.size)
    X = HostCSR(rpt, cols, vals, M, int(ns.max()) if count else 0)
    return X, row_off


def uneven_batch(count: int = 4096, seed: int = 0, nmin: int = 64, nmax: int = 4096,
                 lam: float = 8.0) -> list[HostCSR]:
    """The matrices of uneven_batch_concat as separate CSRs."""
    X, off = uneven_batch_concat(count, seed, nmin, nmax, lam)
    out = []
    for b in range(count):
        r0, r1 = int(off[b]), int(off[b + 1])
        m = X.row_slice(r0, r1)
        m.ncols = r1 - r0
        out.append(m)
    return out
