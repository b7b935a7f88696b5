"""The FEM Hessian generator's oracle (oracle/fem.py) pinned on the CPU:
rest-state closed form, the Jacobi projection against numpy.linalg.eigh,
central differences of the energy, the stress and the pulled-back gradient,
and the rest-state mesh against the synthetic generator's assembled blocks.
Tolerances per test."""
import numpy as np

from hsc_reimpl_amd import systems
from oracle import assembly, fem

MU, LAM = 3.448e7, 3.103e8


def _deformed_tet(rng, amp=0.3):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float) * 0.02
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    x = X + amp * 0.02 * rng.normal(size=(4, 3))
    return x, np.linalg.inv(Dm), abs(np.linalg.det(Dm)) / 6


def test_rest_state_is_the_closed_form_tangent():
    """F = I: mu (d_ij d_ab + d_ib d_ja) + lambda d_ia d_jb, which is also the
    synthetic generator's Neo-Hookean rest tangent (to its eigh rounding)."""
    H = fem.raw_hessian(MU, LAM, np.eye(3))
    d = np.eye(3)
    C = (MU * np.einsum("ij,ab->iajb", d, d) + MU * np.einsum("ib,ja->iajb", d, d)
         + LAM * np.einsum("ia,jb->iajb", d, d))
    np.testing.assert_array_equal(H.reshape(3, 3, 3, 3, order="F").transpose(0, 1, 2, 3),
                                  C.transpose(0, 1, 2, 3).reshape(3, 3, 3, 3))
    E, nu = 1e8, 0.45
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    Hs = systems.neo_hookean_rest_tangent(E, nu).reshape(3, 3, 3, 3)  # [i, a, j, b] row-major vec
    Hf = fem.raw_hessian(mu, lam, np.eye(3)).reshape(3, 3, 3, 3, order="F")  # [i, a, j, b]
    np.testing.assert_allclose(Hf, Hs, rtol=0, atol=1e-9 * np.abs(Hs).max())


def test_projection_matches_eigh_clamp():
    """Compressed and sheared F (indefinite Hessians): the Jacobi projection
    equals Q max(w, 0) Q^T from numpy.linalg.eigh within 1e-10 |H|."""
    rng = np.random.default_rng(0)
    n_indef = 0
    for _ in range(20):
        F = np.eye(3) + 0.6 * rng.normal(size=(3, 3))
        if np.linalg.det(F) <= 0.05:
            continue
        H = fem.raw_hessian(MU, LAM, F)
        w, Q = np.linalg.eigh(0.5 * (H + H.T))
        n_indef += w.min() < 0
        ref = (Q * np.maximum(w, 0)) @ Q.T
        np.testing.assert_allclose(fem.project_spd(H), ref, rtol=0, atol=1e-10 * np.linalg.norm(H))
    assert n_indef > 0


def test_derivatives_by_central_differences():
    """Energy -> first Piola stress -> 9x9 Hessian (vec(F) column-major) ->
    pulled-back tet blocks, each against central differences of the previous
    (relative 1e-6)."""
    rng = np.random.default_rng(1)
    F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    h = 1e-6
    P = fem.first_piola(MU, LAM, F)
    Pn = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            dF = np.zeros((3, 3))
            dF[i, j] = h
            Pn[i, j] = (fem.energy(MU, LAM, F + dF) - fem.energy(MU, LAM, F - dF)) / (2 * h)
    np.testing.assert_allclose(Pn, P, rtol=1e-6, atol=1e-6 * np.abs(P).max())
    H = fem.raw_hessian(MU, LAM, F)
    Hn = np.zeros((9, 9))
    for j in range(3):
        for b in range(3):
            dF = np.zeros((3, 3))
            dF[j, b] = h
            dP = (fem.first_piola(MU, LAM, F + dF) - fem.first_piola(MU, LAM, F - dF)) / (2 * h)
            Hn[:, j + 3 * b] = dP.ravel(order="F")
    np.testing.assert_allclose(Hn, H, rtol=1e-6, atol=1e-6 * np.abs(H).max())
    x, Dinv, vol = _deformed_tet(rng)
    blocks = fem.tet_blocks(x, Dinv, MU, LAM, vol, project=False)
    hx = 1e-9
    for a, b, B in blocks:
        Bn = np.zeros((3, 3))
        for l in range(3):
            xp, xm = x.copy(), x.copy()
            xp[b, l] += hx
            xm[b, l] -= hx
            Bn[:, l] = (fem.tet_gradient(xp, Dinv, MU, LAM, vol)[a] - fem.tet_gradient(xm, Dinv, MU, LAM, vol)[a]) / (2 * hx)
        np.testing.assert_allclose(Bn, B, rtol=1e-5, atol=1e-5 * np.abs(B).max())


def test_rest_mesh_assembles_to_the_generator_blocks():
    """The 2,500-tet cube mesh at rest: oracle tet blocks, assembled, equal the
    synthetic generator's fem_blocks (upper triangle) within 1e-12 relative."""
    nodes, tets = systems.box_mesh(10, 10, 5, 0.02)
    E, nu = 1e8, 0.45
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    R, C, V = [], [], []
    for tet in tets:
        X = nodes[tet]
        Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
        for a, b, B in fem.tet_blocks(X, np.linalg.inv(Dm), mu, lam, abs(np.linalg.det(Dm)) / 6):
            l, r = (a, b) if tet[a] <= tet[b] else (b, a)
            R.append(tet[l])
            C.append(tet[r])
            V.append((B if l == a else B.T).ravel(order="F"))
    r, c, v = assembly.assemble_bcoo(np.array(R), np.array(C), np.array(V))
    I, J, B = systems.fem_blocks(nodes, tets, E, nu)
    up = I <= J
    np.testing.assert_array_equal(r, I[up])
    np.testing.assert_array_equal(c, J[up])
    ref = B[up].transpose(0, 2, 1).reshape(-1, 9)     # row-major blocks -> column-major
    np.testing.assert_allclose(v, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
