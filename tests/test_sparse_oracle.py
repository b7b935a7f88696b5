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
