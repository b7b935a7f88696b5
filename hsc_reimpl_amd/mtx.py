"""MatrixMarket files for replaying linear systems through the C ABI.

A system is dumped the way the libuipc global linear system dumps its
Newton systems (`_dump_A_b`, global_linear_system.cu:39-56 of the
reference): A as a coordinate matrix of the stored 3x3 blocks expanded to
scalar entries (1-based), b as a dense array.  libuipc's own writer
(`utils/matrix_market.h`) is not part of the reference sources, so the
header details follow the MatrixMarket format itself; the block structure is
recovered on reading by grouping scalar entries into 3x3 blocks.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def write_bcoo(path, rows, cols, vals, n_block_rows: int) -> None:
    """3x3-block triplets (vals (n, 9) column-major) as a coordinate real
    general matrix of N = 3 * n_block_rows with 9 entries per block."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    B = np.asarray(vals, dtype=np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)  # row-major
    ii = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2).ravel() + 1
    jj = (3 * cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1).ravel() + 1
    N = 3 * n_block_rows
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{N} {N} {len(ii)}\n")
        np.savetxt(f, np.column_stack([ii, jj, B.ravel()]), fmt=["%d", "%d", "%.17g"])


def read_bcoo(path):
    """Back to (rows, cols, vals (n, 9) column-major, n_block_rows), blocks in
    file order (scalar entries of one block may come in any order; entries of
    a block missing from the file are zero)."""
    with open(path) as f:
        header = f.readline()
        if not header.startswith("%%MatrixMarket matrix coordinate real"):
            raise ValueError("not a coordinate real MatrixMarket matrix")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        N, M, nnz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2) if nnz else np.zeros((0, 3))
    if N != M or N % 3:
        raise ValueError("expects a square matrix of 3x3 blocks")
    i = data[:, 0].astype(np.int64) - 1
    j = data[:, 1].astype(np.int64) - 1
    key = (i // 3) * (N // 3) + (j // 3)
    uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")        # blocks in order of first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    vals = np.zeros((len(uk), 9))
    vals[rank[inv], (i % 3) + 3 * (j % 3)] = data[:, 2]
    uk = uk[order]
    return (uk // (N // 3)).astype(np.int32), (uk % (N // 3)).astype(np.int32), vals, N // 3


def write_vector(path, b) -> None:
    b = np.asarray(b, dtype=np.float64).ravel()
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{len(b)} 1\n")
        np.savetxt(f, b, fmt="%.17g")


def read_vector(path) -> np.ndarray:
    with open(path) as f:
        if not f.readline().startswith("%%MatrixMarket matrix array real"):
            raise ValueError("not an array real MatrixMarket vector")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, m = (int(t) for t in line.split())
        return np.loadtxt(f, ndmin=1).reshape(n * m)
