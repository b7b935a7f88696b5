"""CPU restatement of the ABD-FEM coupling assembly before the HSC path
(ABDFEMLinearSubsystem_assemble_kernel, abd_fem_linear_subsystem.cu:18-61 of
the reference): each contact block H (3x3; rows on an ABD vertex, columns on
an FEM vertex) becomes J^T H (12x3), with J = dx/dq the affine Jacobian at
the ABD vertex's rest position xbar, written as 4 block triplets
(abd_offset + 4 body + k, fem_offset + fem vertex), zeroed when the body or
the FEM vertex is fixed.

The affine map is x = p + A xbar with q = (p, rows of A) -- libuipc's
ABDJacobi, which is not part of the reference sources; the DOF order is the
one the reference's gravity setup uses for J^T g
(affine_body_dynamics.cu:609-613: segment 3 + 3k = g[k] xbar).

Test infrastructure only (the checker of hsc_abd_fem_coupling).  Pinned by
(a) the gravity layout above, (b) J probed from the position map x(q) by
unit vectors, and (c) the assembled coupling equal to the q-x_f block of
G^T H_x G for a global map G over many contacts (tests/test_coupling_oracle.py).
"""
from __future__ import annotations

import numpy as np


def abd_jacobian(xbar):
    """3 x 12 dx/dq at rest position xbar (x = p + A xbar, q = (p, A rows))."""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for i in range(3):
        J[i, 3 + 3 * i:6 + 3 * i] = xbar
    return J


def coupling_blocks(abd_vertex, fem_vertex, hess, v2b, xbar, body_fixed, vertex_fixed, abd_offset=0,
                    fem_offset=0):
    """hess: (n, 3, 3) row-major contact blocks.  Returns (rows, cols,
    vals (4n, 9) column-major) in the kernel's order (contact-major, k minor)."""
    n = len(abd_vertex)
    rows = np.empty(4 * n, dtype=np.int64)
    cols = np.empty(4 * n, dtype=np.int64)
    vals = np.empty((4 * n, 9))
    for c in range(n):
        av, fv = int(abd_vertex[c]), int(fem_vertex[c])
        body = int(v2b[av])
        H12 = abd_jacobian(np.asarray(xbar[av], dtype=np.float64)).T @ np.asarray(hess[c], dtype=np.float64)
        if body_fixed[body] or vertex_fixed[fv]:
            H12 = np.zeros((12, 3))
        for k in range(4):
            rows[4 * c + k] = abd_offset + 4 * body + k
            cols[4 * c + k] = fem_offset + fv
            vals[4 * c + k] = H12[3 * k:3 * k + 3].T.ravel()  # column-major 3x3
    return rows, cols, vals
