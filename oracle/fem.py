"""CPU restatement of the FEM Hessian generator before the HSC path: the
per-tet log-Neo-Hookean Hessian with the SPD projection by bounded cyclic
Jacobi, pulled back to 3x3 node blocks (the reference's overlays:
log_neo_hookean_function.h:93-150, log_neo_hookean_3d.cu:62-135).  libuipc's
fem helpers (F, shape gradients, the Hessian pullback) are not part of the
reference sources; they are restated from their definitions: F = Ds Dm^-1
with Ds = [x1-x0, x2-x0, x3-x0], vec(F) column-major, grad N_a = row a-1 of
Dm^-1 (a > 0), grad N_0 = -(sum).

Test infrastructure only (the checker of hsc_fem_hessian).  Pinned by the
closed-form rest-state tangent, by numpy.linalg.eigh for the projection and
by finite differences of the energy and of the pulled-back gradient.
"""
from __future__ import annotations

import numpy as np


def deformation_gradient(x, Dm_inv):
    """x: (4, 3) node positions; Dm_inv: (3, 3). F = Ds Dm^-1."""
    Ds = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
    return Ds @ Dm_inv


def energy(mu, lam, F):
    J = np.linalg.det(F)
    return mu / 2 * (np.sum(F * F) - 3) - mu * np.log(J) + lam / 2 * np.log(J) ** 2


def first_piola(mu, lam, F):
    J = np.linalg.det(F)
    C = J * np.linalg.inv(F).T                    # cofactor
    return mu * F + (lam * np.log(J) - mu) / J * C


def raw_hessian(mu, lam, F):
    """9x9 d2Psi/dF2 with vec(F) column-major: H[i + 3a, j + 3b]."""
    J = np.linalg.det(F)
    G = np.linalg.inv(F).T                        # cofactor / J
    q = lam * np.log(J) - mu
    H = np.zeros((9, 9))
    for a in range(3):
        for i in range(3):
            for b in range(3):
                for j in range(3):
                    H[i + 3 * a, j + 3 * b] = (mu * (i == j and a == b) + lam * G[i, a] * G[j, b]
                                               - q * G[i, b] * G[j, a])
    return H


def project_spd(H):
    """Bounded cyclic Jacobi eigendecomposition (32 sweeps, threshold
    1e-13 |H|_F), eigenvalues clamped at zero, Q diag(w) Q^T."""
    H = np.array(H, dtype=np.float64)
    Q = np.eye(9)
    threshold = 1e-13 * np.linalg.norm(H)
    for _ in range(32):
        maximum = 0.0
        for p in range(8):
            for q in range(p + 1, 9):
                apq = H[p, q]
                maximum = max(maximum, abs(apq))
                if abs(apq) <= threshold:
                    continue
                tau = (H[q, q] - H[p, p]) / (2 * apq)
                t = np.copysign(1 / (abs(tau) + np.sqrt(1 + tau * tau)), tau)
                c = 1 / np.sqrt(1 + t * t)
                s = t * c
                H[p, p] -= t * apq
                H[q, q] += t * apq
                H[p, q] = H[q, p] = 0.0
                for k in range(9):
                    if k != p and k != q:
                        kp, kq = H[k, p], H[k, q]
                        H[k, p] = H[p, k] = c * kp - s * kq
                        H[k, q] = H[q, k] = s * kp + c * kq
                    qp, qk = Q[k, p], Q[k, q]
                    Q[k, p] = c * qp - s * qk
                    Q[k, q] = s * qp + c * qk
        if maximum <= threshold:
            break
    w = np.maximum(np.diag(H), 0.0)
    return (Q * w) @ Q.T


def shape_gradients(Dm_inv):
    g = np.empty((4, 3))
    g[1:] = Dm_inv
    g[0] = -Dm_inv.sum(axis=0)
    return g


def tet_blocks(x, Dm_inv, mu, lam, vol, dt=1.0, project=True):
    """The 10 node blocks (a <= b by node index) of one tet: list of
    (local a, local b, 3x3 block) with block(k, l) = sum_{p,m} gN_a[p] gN_b[m]
    H[k + 3p, l + 3m] * vol * dt^2."""
    F = deformation_gradient(x, Dm_inv)
    H = raw_hessian(mu, lam, F)
    if project:
        H = project_spd(H)
    H = H * vol * dt * dt
    g = shape_gradients(Dm_inv)
    Hr = H.reshape(3, 3, 3, 3, order="F")          # Hr[k, p, l, m] = H[k + 3p, l + 3m]
    return [(a, b, np.einsum("p,kplm,m->kl", g[a], Hr, g[b])) for a in range(4) for b in range(a, 4)]


def tet_gradient(x, Dm_inv, mu, lam, vol, dt=1.0):
    """(4, 3) nodal gradient: G_a[k] = sum_p P(k, p) gN_a[p] * vol * dt^2."""
    P = first_piola(mu, lam, deformation_gradient(x, Dm_inv))
    return shape_gradients(Dm_inv) @ P.T * vol * dt * dt
