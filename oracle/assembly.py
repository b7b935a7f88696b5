"""CPU restatement of the assembly before the HSC path: general 3x3-block
triplets -> the upper-triangle BCOO of unique blocks the plugin receives
(libuipc's MatrixConverter::ge2sym + convert, called at
global_linear_system.cu:232-234 of the reference; MatrixConverter is not
part of the reference sources, so the semantics follow what the plugin's
split assumes of its input: one stored block per unordered pair with
row <= col, diagonal blocks once, linear_hsc.cu:211-247).

Test infrastructure only (the checker of hsc_assemble_bcoo).  Duplicates are
summed in input order, a lower block is folded onto (col, row) transposed or
dropped.  Pinned by an independent dictionary accumulation and by the
round trip from a generated system's upper BCOO.
"""
from __future__ import annotations

import numpy as np


def assemble_bcoo(rows, cols, vals, fold_lower=True):
    """rows, cols: int block indices; vals: (n, 9) column-major blocks.
    Returns (rows, cols, vals) of the unique upper blocks sorted by (row, col)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64).reshape(-1, 9)
    lower = rows > cols
    keep = ~lower | bool(fold_lower)
    r = np.where(lower, cols, rows)[keep]
    c = np.where(lower, rows, cols)[keep]
    v = vals.copy()
    # a folded lower block lands transposed: column-major v[r + 3c] -> (c, r)
    v[lower] = v[lower].reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)
    v = v[keep]
    key = (r << 32) | c
    order = np.argsort(key, kind="stable")          # equal keys stay in input order
    key, v = key[order], v[order]
    heads = np.ones(len(key), dtype=bool)
    heads[1:] = key[1:] != key[:-1]
    seg = np.cumsum(heads) - 1
    out = np.zeros((int(heads.sum()), 9))
    for q in range(len(key)):                      # sequential, in input order within a block
        out[seg[q]] += v[q]
    uk = key[heads]
    return (uk >> 32).astype(np.int32), (uk & 0xFFFFFFFF).astype(np.int32), out
