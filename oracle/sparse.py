"""Sparse restatement of the HSC solve for systems too large for dense form.

Same algorithm as oracle.hsc_core (core.py:22-168 of the reference) but with
D held as a SciPy sparse matrix, plus the plugin's setup split of the BCOO
input (linear_hsc.cu:195-247 of the reference).  Test oracle / CPU baseline
only -- "sparse restatement, not core.py".
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from . import hsc_core


def split_bcoo(rows, cols, vals, qo, qm, fo, fn):
    """Split the stored-triangle BCOO into D (FEM, sparse), A (ABD, dense),
    E (qm x active, dense) and the sorted active FEM DOF list.

    Follows linear_hsc.cu:211-247: one triangle stored, diagonal blocks not
    doubled, off-diagonal blocks mirrored/transposed; an FEM node is active
    when any coupling block touching it has nonzero Frobenius norm, and then
    all three of its DOFs are active; E columns follow the sorted active list.
    """
    rows = np.asarray(rows, dtype=np.int64) * 3
    cols = np.asarray(cols, dtype=np.int64) * 3
    B = np.asarray(vals, dtype=np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)  # row-major
    iq = (rows >= qo) & (rows < qo + qm)
    jq = (cols >= qo) & (cols < qo + qm)
    ifem = (rows >= fo) & (rows < fo + fn)
    jfem = (cols >= fo) & (cols < fo + fn)
    # D (full symmetric)
    ff = ifem & jfem
    r, c, b = rows[ff] - fo, cols[ff] - fo, B[ff]
    off = r != c
    ii = (r[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2)
    jj = (c[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1)
    data = np.concatenate([b.ravel(), b[off].transpose(0, 2, 1).ravel()])
    ri = np.concatenate([ii.ravel(), jj[off].transpose(0, 2, 1).ravel()])
    ci = np.concatenate([jj.ravel(), ii[off].transpose(0, 2, 1).ravel()])
    D = sp.csr_matrix((data, (ri, ci)), shape=(fn, fn))
    # A (dense)
    A = np.zeros((qm, qm))
    for rr, cc, bb in zip(rows[iq & jq] - qo, cols[iq & jq] - qo, B[iq & jq]):
        A[rr:rr + 3, cc:cc + 3] += bb
        if rr != cc:
            A[cc:cc + 3, rr:rr + 3] += bb.T
    # E (dense over the active columns)
    cpl = (iq ^ jq)
    nz = cpl & (np.einsum("kij,kij->k", B, B) != 0)
    fnode = np.where(iq, cols, rows)[nz]
    if np.any((fnode < fo) | (fnode >= fo + fn)):
        raise ValueError("Unexpected HSC subsystem coupling")
    active = np.unique((fnode[:, None] + np.arange(3)[None, :]).ravel())
    E = np.zeros((qm, len(active)))
    for q_is_row, rr, cc, bb in zip(iq[nz], rows[nz], cols[nz], B[nz]):
        q = (rr if q_is_row else cc) - qo
        f = cc if q_is_row else rr
        k = int(np.searchsorted(active, f))
        E[q:q + 3, k:k + 3] += bb if q_is_row else bb.T
    return D, A, E, active - fo


class BlockJacobi3:
    """w * blockdiag(D)^-1 applied per 3x3 node block."""

    def __init__(self, D, omega):
        n = D.shape[0]
        Dd = D.tocsr()
        blocks = np.zeros((n // 3, 3, 3))
        for a in range(3):
            for c in range(3):
                blocks[:, a, c] = Dd[np.arange(a, n, 3), :][:, np.arange(c, n, 3)].diagonal()
        for blk in blocks:
            np.linalg.cholesky(blk)
        self.inv = np.linalg.inv(blocks)
        self.omega = omega

    def __call__(self, v):
        return self.omega * np.einsum("kij,kj->ki", self.inv, v.reshape(-1, 3)).ravel()


class SparseNeumann:
    """Neumann(D, 3, order, omega) with sparse D (core.py:74-92)."""

    def __init__(self, D, order=5, omega=0.7):
        self.D, self.order, self.omega = D, order, omega
        self.J = BlockJacobi3(D, omega)

    def __call__(self, rhs):
        t = self.J(rhs)
        z = t.copy()
        for _ in range(self.order - 1):
            t = t - self.J(self.D @ t)
            z += t
        return z


class SparsePolyJacobi:
    """Block-Jacobi polynomial preconditioner with per-pass weights (the
    build's Chebyshev option; no reference counterpart):
    z0 = w0 B^-1 r, z_j = z_(j-1) + w_j B^-1 (r - D z_(j-1)).  With every w_j
    equal to omega this is SparseNeumann's series in Horner form."""

    def __init__(self, D, omegas):
        self.D, self.omegas = D, list(omegas)
        self.J = BlockJacobi3(D, 1.0)

    def __call__(self, r):
        z = self.omegas[0] * self.J(r)
        for w in self.omegas[1:]:
            z = z + w * self.J(r - self.D @ z)
        return z


def leja_order(points):
    """Leja order of `points`: the largest first, then repeatedly the point
    maximising the product of its distances to those already taken (sum of
    logs, candidates scanned in ascending order, first maximum kept)."""
    import math
    rest = sorted(points)
    out = [rest.pop()]
    while rest:
        best, best_v = 0, -math.inf
        for i, x in enumerate(rest):
            v = 0.0
            for o in out:
                v += math.log(abs(x - o))
            if v > best_v:
                best, best_v = i, v
        out.append(rest.pop(best))
    return out


def chebyshev_weights(lam_est, m, ratio, safety=1.2):
    """Per-pass weights 1/mu_j, mu_j the roots of the degree-m Chebyshev
    polynomial on [b/ratio, b], b = safety * lam_est, in Leja order (keeps
    every partial product of the passes' factors, and every remaining one,
    bounded, so that rounding in FP32 passes is not amplified)."""
    b = safety * lam_est
    a = b / ratio
    mu = leja_order([0.5 * (b + a) + 0.5 * (b - a) * np.cos((2 * j + 1) * np.pi / (2 * m)) for j in range(m)])
    return [1.0 / x for x in mu], a, b


class SparseHSC:
    """HSC with sparse D (core.py:141-168).  E is the compacted qm x active
    coupling.  The ABD inverse is either exact FP64 (Cholesky; core.py's
    exact_affine=True) or, with abd_inverse="neumann", the drop-in plugin's
    block Neumann polynomial of the same order and omega as the FEM passes,
    doubled while ||I - A P||_F / sqrt(qm) > abd_tol when `adaptive`
    (linear_hsc.cu:263-297, hsc_core.neumann_adaptive)."""

    def __init__(self, D, A, E, active, order=5, omega=0.7, aca_rtol=None, aca_max_rank=None,
                 abd_inverse="exact", adaptive=True, abd_tol=1e-5):
        self.D, self.A, self.E, self.active = D, A, E, np.asarray(active)
        self.n = D.shape[0]
        self.factors = hsc_core.aca(E, rtol=aca_rtol, max_rank=aca_max_rank)
        self.U, self.Vc = self.factors.U, self.factors.V
        if abd_inverse == "exact":
            self.cho, self.P, self.doublings, self.abd_residual = sla.cho_factor(A), None, 0, 0.0
        elif abd_inverse == "neumann":
            self.cho = None
            self.P, self.doublings, self.abd_residual = hsc_core.neumann_adaptive(
                np.asarray(A), order=order, omega=omega, tol=abd_tol, adaptive=adaptive)
        else:
            raise ValueError(abd_inverse)
        self.precond_op = SparseNeumann(D, order, omega)
        C = self.U.T @ self.Ainv(self.U)
        self.C = 0.5 * (C + C.T)

    @classmethod
    def from_system(cls, sysm, **kw):
        D, A, E, active = split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm,
                                     sysm.fo, sysm.fn)
        return cls(D, A, E, active, **kw)

    def precond(self, r):
        return self.precond_op(r)

    def Ainv(self, v):
        return sla.cho_solve(self.cho, v) if self.P is None else self.P @ v

    def Vt(self, p):
        return self.Vc.T @ p[self.active]

    def V_apply(self, c):
        out = np.zeros(self.n)
        out[self.active] = self.Vc @ c
        return out

    def matvec(self, p):
        return self.D @ p - self.V_apply(self.C @ self.Vt(p))

    def reduced_rhs(self, bx, bq):
        return bx - self.V_apply(self.U.T @ self.Ainv(bq))

    def back_substitute(self, bq, dx):
        return self.Ainv(bq - self.U @ self.Vt(dx))

    def solve(self, bx, bq, rtol=1e-8, maxiter=20000, dropin=True):
        rhs = self.reduced_rhs(bx, bq)
        if dropin:
            dx, it, rel = hsc_core.pcg_dropin(self.matvec, rhs, self.precond_op, rtol, maxiter)
            stats = {"iterations": it, "relative_residual": rel, "converged": rel <= rtol}
        else:
            dx, stats = hsc_core.pcg(self.matvec, rhs, self.precond_op, rtol, maxiter)
        return dx, self.back_substitute(bq, dx), stats


def direct_solve(sysm):
    """FP64 sparse direct solve of the full system H x = b (SuperLU)."""
    import scipy.sparse.linalg as spla
    H = sysm.sparse().tocsc()
    return spla.spsolve(H, sysm.b)
