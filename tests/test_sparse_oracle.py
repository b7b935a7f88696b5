"""The sparse restatement (used at full size and as the CPU baseline) agrees
with the dense oracle and with SciPy's sparse direct solver."""
import numpy as np
import pytest

from hsc_reimpl_amd import systems
from oracle import hsc_core, sparse as osp
from tests.conftest import load_fixture


@pytest.mark.parametrize("stiffness", [1e2, 1e6])
def test_split_and_solve_match_dense_oracle(stiffness):
    f = load_fixture(stiffness)
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    D, A, E, act = osp.split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    np.testing.assert_array_equal(D.toarray(), f["D"])
    np.testing.assert_allclose(A, f["A"], rtol=0, atol=1e-13 * np.abs(f["A"]).max())  # fixture A is symmetric to round-off
    np.testing.assert_array_equal(E, f["E"][:, act])
    assert np.all(np.abs(f["E"][:, np.setdiff1d(np.arange(len(f["D"])), act)]) == 0)
    dense = hsc_core.HSC(f["D"], f["A"], f["E"], exact_affine=True)
    sp_ = osp.SparseHSC(D, A, E, act)
    n = len(f["D"])
    rng = np.random.default_rng(3)
    v = rng.normal(size=n)
    np.testing.assert_allclose(sp_.matvec(v), dense.matvec(v), rtol=1e-11, atol=1e-11 * np.abs(dense.matvec(v)).max())
    np.testing.assert_allclose(sp_.precond(v), dense.deformable_preconditioner(v), rtol=1e-11, atol=1e-14)
    dx, dq, st = sp_.solve(f["b"][:n], f["b"][n:], rtol=1e-10, maxiter=1000, dropin=False)
    x, st2 = dense.solve(f["b"], rtol=1e-10)
    assert st["iterations"] == st2["iterations"]
    np.testing.assert_allclose(np.concatenate([dx, dq]), x, rtol=1e-8, atol=1e-12)


def test_cube_system_direct_solve():
    sysm = systems.cube_system(seed=0)
    xd = osp.direct_solve(sysm)
    H = sysm.dense()
    np.testing.assert_allclose(H @ xd, sysm.b, atol=1e-9 * np.abs(sysm.b).max())
    ref = osp.SparseHSC.from_system(sysm, order=2, omega=0.6, aca_rtol=1e-6)
    n = sysm.fn
    dx, dq, st = ref.solve(sysm.b[:n], sysm.b[n:], rtol=1e-8, maxiter=5000)
    assert st["converged"]
    x = np.concatenate([dx, dq])
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-7


def test_poly_jacobi_with_equal_weights_is_neumann():
    """The per-pass-weight preconditioner with every weight = omega is the
    Neumann series (whose oracle is pinned by the reference's histories)."""
    sysm = systems.cube_system(seed=0)
    D, _, _, _ = osp.split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    r = np.random.default_rng(3).normal(size=sysm.fn)
    for order, omega in ((2, 0.6), (5, 0.7)):
        a = osp.SparseNeumann(D, order, omega)(r)
        b = osp.SparsePolyJacobi(D, [omega] * order)(r)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-14 * np.abs(a).max())


def test_poly_jacobi_residual_is_the_product_of_its_factors():
    """Unequal weights: the passes' residual r - D M(r) equals
    prod_j (I - w_j D B^-1) r, applied factor by factor (B = D's 3x3 diagonal
    blocks), i.e. the recursion uses each weight exactly once."""
    sysm = systems.cube_system(seed=0)
    D, _, _, _ = osp.split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    Dd = D.tocsr()
    n = Dd.shape[0]
    Binv = np.zeros((n // 3, 3, 3))
    for i in range(n // 3):
        Binv[i] = np.linalg.inv(Dd[3 * i:3 * i + 3, 3 * i:3 * i + 3].toarray())
    apply_binv = lambda v: np.einsum("nij,nj->ni", Binv, v.reshape(-1, 3)).ravel()
    r = np.random.default_rng(5).normal(size=n)
    w, _, _ = osp.chebyshev_weights(2.9, 6, 50.0)
    z = osp.SparsePolyJacobi(D, w)(r)
    e = r.copy()
    for wj in w:
        e = e - wj * (Dd @ apply_binv(e))
    res = r - Dd @ z
    np.testing.assert_allclose(res, e, rtol=1e-9, atol=1e-11 * np.abs(r).max())


def test_chebyshev_weights_leja_order_bounds_partial_products():
    """The weights are the reciprocal Chebyshev roots (a permutation of the
    ascending set) whose partial products prod_{j<k} (1 - w_j mu) and
    remainders prod_{j>=k} stay below 1e3 on [0, b] up to degree 32, where
    ascending roots reach 1e11 (the amplification FP32 passes suffer)."""
    for m, ratio in ((10, 300.0), (24, 1000.0), (32, 3000.0)):
        w, a, b = osp.chebyshev_weights(2.9, m, ratio)
        mu = sorted(0.5 * (b + a) + 0.5 * (b - a) * np.cos((2 * j + 1) * np.pi / (2 * m)) for j in range(m))
        np.testing.assert_allclose(sorted(1.0 / np.array(w)), mu, rtol=1e-14)
        assert np.isclose(1.0 / w[0], max(mu), rtol=1e-14)
        x = np.linspace(0.0, b, 4001)

        def worst(order):
            pre = suf = 1.0
            for k in range(m + 1):
                pre = max(pre, np.abs(np.prod([1 - x / v for v in order[:k]], axis=0)).max())
                suf = max(suf, np.abs(np.prod([1 - x / v for v in order[k:]], axis=0)).max())
            return max(pre, suf)

        assert worst([1.0 / v for v in w]) < 1e3
        if m >= 24:
            assert worst(mu) > 1e10


def test_chebyshev_weights_give_the_chebyshev_residual_polynomial():
    """For D = diag(mu) with B = I the residual polynomial of the passes,
    1 - mu q(mu) = prod_j (1 - w_j mu), equals T_m((b+a-2mu)/(b-a)) /
    T_m((b+a)/(b-a)) (numpy's Chebyshev series), whose max on [a, b] is
    1 / T_m((b+a)/(b-a)); and it stays in (-1, 1) on (0, b]."""
    from numpy.polynomial import chebyshev as Ch
    for m, ratio in ((5, 100.0), (2, 10.0), (8, 30.0)):
        w, a, b = osp.chebyshev_weights(2.5, m, ratio)
        mu = np.linspace(1e-6, b, 4001)
        res = np.ones_like(mu)
        for wj in w:
            res *= 1.0 - wj * mu
        x0 = (b + a) / (b - a)
        Tm = lambda x: Ch.chebval(x, [0] * m + [1])
        ref = Tm((b + a - 2 * mu) / (b - a)) / Tm(x0)
        inside = mu >= a
        np.testing.assert_allclose(res[inside], ref[inside], rtol=1e-9, atol=1e-12)
        assert np.abs(res[inside]).max() <= 1.0 / Tm(x0) * (1 + 1e-9)
        assert np.all(res < 1.0) and np.all(res > -1.0)
