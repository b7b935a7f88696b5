"""Dense NumPy restatement of the reference HSC algorithm (test oracle only).

Follows reproduction/core.py of the reference.  The arithmetic order of each
step is kept where it decides the floating-point result (the saved residual
histories in experiments.json are compared to 1e-9 relative), the code itself
is written independently.

Citations are to reproduction/core.py unless stated otherwise.
"""
from __future__ import annotations

import numpy as np

_EPS = np.finfo(np.float64).eps
_TINY = np.finfo(np.float64).tiny


class ACAFactors:
    """Result of full-pivot ACA: E ~= U @ V.T (core.py:10-19)."""

    def __init__(self, U, V, residuals, pivots):
        self.U = U
        self.V = V
        self.residuals = residuals
        self.pivots = pivots

    @property
    def rank(self) -> int:
        return int(self.U.shape[1])


def aca(E, rtol=None, max_rank=None) -> ACAFactors:
    """Full-pivot adaptive cross approximation (core.py:22-50).

    Pivot = first maximum of |R| in row-major order (np.argmax semantics).
    Stops at the rank limit, when ||R||_F/||E||_F <= rtol (0 when rtol is
    None), or when the pivot is at round-off level eps*max(||E||_F, tiny).
    """
    R = np.array(E, dtype=np.float64, copy=True)
    if R.ndim != 2 or not np.all(np.isfinite(R)) or (rtol is not None and rtol < 0):
        raise ValueError("ACA requires a finite matrix and nonnegative tolerance")
    m, n = R.shape
    limit = min(m, n) if max_rank is None else min(max_rank, m, n)
    scale = float(np.linalg.norm(R))
    stop = 0.0 if rtol is None else float(rtol)
    history = [1.0 if scale else 0.0]
    pivots, cols_u, cols_v = [], [], []
    while len(cols_u) < limit and history[-1] > stop:
        flat = int(np.argmax(np.abs(R)))
        i, j = divmod(flat, n)
        piv = R[i, j]
        if abs(piv) <= _EPS * max(scale, _TINY):
            break
        u = R[:, j].copy()
        v = R[i, :].copy() / piv
        cols_u.append(u)
        cols_v.append(v)
        pivots.append([i, j, float(piv)])
        R -= np.outer(u, v)
        history.append(float(np.linalg.norm(R) / scale))
    U = np.column_stack(cols_u) if cols_u else np.zeros((m, 0))
    V = np.column_stack(cols_v) if cols_v else np.zeros((n, 0))
    return ACAFactors(U, V, history, pivots)


class BlockInverse:
    """Exact inverses of the diagonal blocks of H (core.py:53-71).

    Each block must be SPD (a Cholesky factorisation is attempted first, as
    the reference does, so a non-SPD block raises LinAlgError).
    """

    def __init__(self, H, sizes):
        self.slices, self.inverses = [], []
        start = 0
        for size in sizes:
            sl = slice(start, start + size)
            blk = H[sl, sl]
            np.linalg.cholesky(blk)
            self.slices.append(sl)
            self.inverses.append(np.linalg.solve(blk, np.eye(size)))
            start += size
        if start != H.shape[0]:
            raise ValueError("Block sizes must cover the matrix")

    def __call__(self, x):
        out = np.empty_like(x, dtype=np.float64)
        for sl, inv in zip(self.slices, self.inverses):
            out[sl] = inv @ x[sl]
        return out


class Neumann:
    """Truncated Neumann series sum_{k<r} (I - w B^-1 H)^k w B^-1 (core.py:74-92)."""

    def __init__(self, H, block_size, order=5, omega=0.7):
        if order < 1 or not (0 < omega <= 1) or H.shape[0] % block_size:
            raise ValueError("Invalid Neumann parameters")
        self.H = H
        self.order = order
        self.omega = omega
        self.Binv = BlockInverse(H, [block_size] * (H.shape[0] // block_size))

    def __call__(self, rhs):
        # t_0 = w B^-1 rhs; t_k = t_{k-1} - w B^-1 (H t_{k-1}); z = sum t_k
        t = self.omega * self.Binv(rhs)
        z = t.copy()
        for _ in range(self.order - 1):
            t = t - self.omega * self.Binv(self.H @ t)
            z += t
        return z

    def spectral_radius(self) -> float:
        T = np.eye(len(self.H)) - self.omega * self.Binv(self.H)
        return float(np.max(np.abs(np.linalg.eigvals(T))))


class AbdInverseError(ArithmeticError):
    """The adaptive ABD inverse missed its tolerance or diverged
    (linear_hsc.cu:292-297 of the reference)."""


def neumann_adaptive(A, order=5, omega=0.7, tol=1e-5, adaptive=True, max_doublings=48, strict=True):
    """ABD inverse of the drop-in plugin (linear_hsc.cu:263-297 of the reference).

    P_r = sum_{k<r} (I - w B^-1 A)^k w B^-1 with B the 12x12 diagonal blocks
    of A (each Cholesky-checked), symmetrised; then, when `adaptive`, while
    ||I - A P||_F / sqrt(qm) > tol and fewer than `max_doublings` doublings
    were made: P <- P + P (I - A P), symmetrised.  Raises AbdInverseError when
    a doubling gives a non-finite P or residual, or (when `strict`) when the
    tolerance is still missed.  Returns (P, doublings, residual).
    """
    qm = A.shape[0]
    if qm % 12:
        raise ValueError("ABD block size must be a multiple of 12")
    Binv = np.zeros_like(A)
    for i in range(0, qm, 12):
        blk = A[i:i + 12, i:i + 12]
        np.linalg.cholesky(blk)
        Binv[i:i + 12, i:i + 12] = np.linalg.solve(blk, np.eye(12))
    eye = np.eye(qm)
    T = eye - omega * Binv @ A
    term = omega * Binv
    P = term.copy()
    for _ in range(1, order):
        term = T @ term
        P += term
    P = 0.5 * (P + P.T)
    R = eye - A @ P
    res = float(np.linalg.norm(R) / np.sqrt(qm))
    doublings = 0
    while adaptive and res > tol and doublings < max_doublings:
        P = P + P @ R
        P = 0.5 * (P + P.T)
        R = eye - A @ P
        res = float(np.linalg.norm(R) / np.sqrt(qm))
        doublings += 1
        if not (np.all(np.isfinite(P)) and np.isfinite(res)):
            raise AbdInverseError(f"ABD Neumann extension diverged at doubling {doublings}")
    if strict and adaptive and res > tol:
        raise AbdInverseError(f"ABD inverse residual {res} exceeds {tol} after {doublings} doublings")
    return P, doublings, res


def pcg(matvec, b, preconditioner=None, rtol=1e-10, maxiter=1000):
    """PCG with the TRUE relative residual tested every iteration (core.py:95-138).

    * stop when ||b - A x|| / ||b|| <= rtol (unsquared);
    * reliable restart (r <- true residual, p <- M r) when the recursive
      residual has converged but the true one has not;
    * explicit non-SPD breakdown (pAp or rz not finite-positive).
    """
    x = np.zeros_like(b)
    nb = np.linalg.norm(b)
    if nb == 0:
        return x, {"iterations": 0, "relative_residual": 0.0,
                   "converged": True, "history": [0.0]}
    M = preconditioner if preconditioner is not None else (lambda v: v.copy())
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = float(r @ z)
    hist = [1.0]
    it = 0
    for it in range(1, maxiter + 1):
        Ap = matvec(p)
        pAp = float(p @ Ap)
        if not np.isfinite(pAp) or pAp <= 0 or rz <= 0:
            raise ArithmeticError(f"PCG non-SPD breakdown at iteration {it}")
        alpha = rz / pAp
        x += alpha * p
        r -= alpha * Ap
        true_r = b - matvec(x)
        rel = float(np.linalg.norm(true_r) / nb)
        hist.append(rel)
        if rel <= rtol:
            break
        if np.linalg.norm(r) <= rtol * nb:
            r = true_r
            z = M(r)
            rz = float(r @ z)
            p = z.copy()
            continue
        z = M(r)
        rz_next = float(r @ z)
        p = z + (rz_next / rz) * p
        rz = rz_next
    return x, {"iterations": it, "relative_residual": hist[-1],
               "converged": hist[-1] <= rtol, "history": hist}


def pcg_dropin(matvec, b, preconditioner, rtol=1e-4, maxiter=2000):
    """Schur PCG with the drop-in plugin's stopping rule (linear_hsc.cu:359-394).

    The recursive residual is tested each iteration; on convergence ONE true
    residual check is made, and if it fails PCG restarts from the true
    residual.  Returns (x, iterations, relative) like the plugin.
    """
    x = np.zeros_like(b)
    nb = float(np.linalg.norm(b))
    rel = 1.0 if nb > 0 else 0.0
    it = 0
    if nb > 0:
        r = b.copy()
        z = preconditioner(r)
        p = z.copy()
        rz = float(r @ z)
        while it < maxiter and rel > rtol:
            Ap = matvec(p)
            pAp = float(p @ Ap)
            if not (np.isfinite(pAp) and pAp > 0 and np.isfinite(rz) and rz > 0):
                raise ArithmeticError(f"HSC Schur PCG breakdown at iteration {it}")
            alpha = rz / pAp
            x += alpha * p
            r -= alpha * Ap
            it += 1
            rel = float(np.linalg.norm(r)) / nb
            if rel <= rtol:
                r = b - matvec(x)
                rel = float(np.linalg.norm(r)) / nb
                if rel <= rtol:
                    break
                z = preconditioner(r)
                p = z.copy()
                rz = float(r @ z)
                continue
            z = preconditioner(r)
            rz_next = float(r @ z)
            p = z + (rz_next / rz) * p
            rz = rz_next
    return x, it, rel


class HSC:
    """Low-rank Schur elimination of the ABD block (core.py:141-168).

    H = [[D, E^T], [E, A]] with FEM first.  E ~= U V^T by ACA; A^-1 is exact
    (exact_affine=True) or the truncated block Neumann series; the Schur
    operator is S p = D p - V (C (V^T p)) with C = U^T A^-1 U symmetrised.
    """

    def __init__(self, D, A, E, order=5, omega=0.7, aca_rtol=None,
                 exact_affine=False, aca_max_rank=None, check_schur=True, affine_P=None):
        self.D, self.A, self.E = D, A, E
        self.factors = aca(E, rtol=aca_rtol, max_rank=aca_max_rank)
        self.U, self.V = self.factors.U, self.factors.V
        self.affine_neumann = Neumann(A, 12, order, omega)
        if affine_P is not None:
            # an explicit A^-1 approximation (the drop-in plugin's adaptive P)
            self.affine_inverse = lambda v: affine_P @ v
        elif exact_affine:
            self.affine_inverse = lambda v: np.linalg.solve(A, v)
        else:
            self.affine_inverse = self.affine_neumann
        self.deformable_preconditioner = Neumann(D, 3, order, omega)
        C = self.U.T @ self.affine_inverse(self.U)
        self.C = (C + C.T) * 0.5
        if check_schur:
            self.schur = D - self.V @ self.C @ self.V.T
            np.linalg.cholesky(self.schur)

    def matvec(self, p):
        return self.D @ p - self.V @ (self.C @ (self.V.T @ p))

    def reduced_rhs(self, b):
        n = len(self.D)
        return b[:n] - self.V @ (self.U.T @ self.affine_inverse(b[n:]))

    def back_substitute(self, b, dx):
        n = len(self.D)
        return self.affine_inverse(b[n:] - self.U @ (self.V.T @ dx))

    def solve(self, b, rtol=1e-10, maxiter=1000):
        rhs = self.reduced_rhs(b)
        dx, stats = pcg(self.matvec, rhs, self.deformable_preconditioner, rtol, maxiter)
        dq = self.back_substitute(b, dx)
        return np.concatenate((dx, dq)), stats


def refinement(H, b, solver, rtol=1e-9, maxsteps=50):
    """Inexact-Newton correction loop on a quadratic (core.py:171-198)."""
    x = np.zeros_like(b)
    nb = np.linalg.norm(b)
    if nb == 0:
        return x, {"steps": 0, "history": [0.0], "converged": True,
                   "inner_iterations": [], "step_lengths": []}
    hist, inner, steps = [1.0], [], []
    for _ in range(maxsteps):
        res = b - H @ x
        d, info = solver.solve(res)
        inner.append(info["iterations"])
        num = float(res @ d)
        den = float(d @ H @ d)
        if num <= 0 or den <= 0:
            raise ArithmeticError("The approximate HSC direction is not descent")
        a = min(1.0, num / den)
        x += a * d
        steps.append(a)
        hist.append(float(np.linalg.norm(b - H @ x) / nb))
        if hist[-1] <= rtol:
            break
    return x, {"steps": len(inner), "history": hist, "converged": hist[-1] <= rtol,
               "inner_iterations": inner, "step_lengths": steps}


def affine_projection(rest):
    """3x12 ABD Jacobian of a point with rest position `rest` (core.py:201-206).

    Row k holds `rest` in columns 3k..3k+2 and 1 in column 9+k.
    """
    J = np.zeros((3, 12))
    for k in range(3):
        J[k, 3 * k:3 * k + 3] = rest
        J[k, 9 + k] = 1.0
    return J


def manufactured_system(stiffness=1e6, seed=2026, nx=10, ny=3, nz=2):
    """Contact-like SPD fixture generator (core.py:209-260).

    The random stream is consumed in exactly the reference's order so the
    committed fixture_*.npz files are reproduced bit for bit.
    Returns D, A, E, H, b, x_star (FEM DOFs first, then 3 ABD bodies).
    """
    rng = np.random.default_rng(seed)
    grid = np.array([(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)],
                    dtype=float)
    nv = len(grid)
    m = 36
    n = 3 * nv
    # graph Laplacian of the unit lattice (manhattan distance 1 neighbours)
    dist = np.abs(grid[:, None, :] - grid[None, :, :]).sum(axis=2)
    adj = (dist == 1).astype(float)
    lap = np.diag(adj.sum(axis=1)) - adj
    mat = np.array([[2.0, 0.2, 0.1], [0.2, 1.0, 0.15], [0.1, 0.15, 1.5]])
    D = np.kron(40 * lap + np.eye(nv), mat)
    A = np.zeros((m, m))
    for body in range(3):
        Q, _ = np.linalg.qr(rng.normal(size=(12, 12)))
        sl = slice(12 * body, 12 * body + 12)
        A[sl, sl] = Q @ np.diag(np.geomspace(10.0, stiffness, 12)) @ Q.T
    H = np.zeros((n + m, n + m))
    H[:n, :n] = D
    H[n:, n:] = A
    weights = np.geomspace(1.0, 1e3, 18)
    for c in range(18):
        node = int(rng.integers(nv))
        body = c % 3
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        rest = rng.uniform(-0.5, 0.5, 3)
        g = np.zeros(n + m)
        g[3 * node:3 * node + 3] = nrm
        g[n + 12 * body:n + 12 * body + 12] = -nrm @ affine_projection(rest)
        H += weights[c] * np.outer(g, g)
    for body in range(2):
        g = np.zeros(n + m)
        g[n + 12 * body + 9] = 1.0
        g[n + 12 * (body + 1) + 9] = -1.0
        H += 2.0 * np.outer(g, g)
    x_star = rng.normal(size=n + m)
    b = H @ x_star
    s = np.linalg.norm(b)
    b, x_star = b / s, x_star / s
    return H[:n, :n].copy(), H[n:, n:].copy(), H[n:, :n].copy(), H, b, x_star
