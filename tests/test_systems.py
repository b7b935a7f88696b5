"""Synthetic workload generator: sizes stated by the workloads, symmetry,
positive definiteness, coupling rank, BCOO format invariants."""
import numpy as np

from hsc_reimpl_amd import systems
from oracle import hsc_core, sparse as osp


def test_cube_system_shape_and_spd():
    s = systems.cube_system(seed=0)
    assert s.info["tets"] == 2500 and s.info["nb"] == 726 and s.fn == 2178 and s.qm == 12
    assert np.all(s.rows <= s.cols)
    keys = s.rows.astype(np.int64) * (s.N // 3) + s.cols
    assert len(np.unique(keys)) == len(keys)
    H = s.dense()
    np.testing.assert_allclose(H, H.T, atol=1e-6 * np.abs(H).max())
    assert np.linalg.eigvalsh(0.5 * (H + H.T))[0] > 0
    D, A, E, act = osp.split_bcoo(s.rows, s.cols, s.vals, s.qo, s.qm, s.fo, s.fn)
    assert len(act) == 192
    f = hsc_core.aca(E, rtol=1e-6)
    assert f.rank == 6


def test_torus_mesh_counts():
    nodes, tets, outer = systems.torus_mesh(0.45, 0.12, 0.06, 70, 16, 2)
    assert len(tets) == 11200 and len(nodes) == 70 * 16 * 3 and len(outer) == 70 * 16
    X = nodes[tets]
    vol = np.einsum("ij,ij->i", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6
    assert np.all(vol > 0)
    # closed surface of a hollow torus: total volume = 2 pi^2 R (r_out^2 - r_in^2) within mesh error
    exact = 2 * np.pi ** 2 * 0.45 * (0.12 ** 2 - 0.06 ** 2)
    assert abs(vol.sum() - exact) / exact < 0.05


def test_box_mesh_volume_and_stiffness_nullspace():
    nodes, tets = systems.box_mesh(3, 2, 2, 0.1)
    I, J, B = systems.fem_blocks(nodes, tets, 1e6, 0.3)
    n = len(nodes)
    K = np.zeros((3 * n, 3 * n))
    for i, j, b in zip(I, J, B):
        K[3 * i:3 * i + 3, 3 * j:3 * j + 3] += b
    np.testing.assert_allclose(K, K.T, atol=1e-6)
    # rigid translations and infinitesimal rotations are in the null space
    t = np.tile([1.0, 0.0, 0.0], n)
    rot = np.cross(np.array([0.0, 0.0, 1.0]), nodes).ravel()
    assert np.abs(K @ t).max() < 1e-6 * np.abs(K).max()
    assert np.abs(K @ rot).max() < 1e-6 * np.abs(K).max()
    w = np.linalg.eigvalsh(K)
    assert w[0] > -1e-6 * w[-1] and np.sum(w < 1e-8 * w[-1]) == 6


def test_from_dense_roundtrip():
    D, A, E, H, b, _ = hsc_core.manufactured_system(1e4, nx=4, ny=2, nz=2)
    s = systems.from_dense(D, A, E, b)
    np.testing.assert_allclose(s.dense(), H, rtol=0, atol=1e-13 * np.abs(H).max())


def test_replicate_fem_is_block_diagonal_with_one_coupling():
    """replicate_fem: the FEM block repeated on the diagonal, the ABD block and
    its coupling kept once (to the first copy), nothing between copies."""
    s = systems.cube_system(seed=0)
    R = systems.replicate_fem(s, 3)
    assert (R.fn, R.qm, R.qo) == (3 * s.fn, s.qm, 3 * s.fn)
    assert np.all(R.rows <= R.cols)
    H, HR, n = s.dense(), R.dense(), s.fn
    for k in range(3):
        np.testing.assert_array_equal(HR[k * n:(k + 1) * n, k * n:(k + 1) * n], H[:n, :n])
    np.testing.assert_array_equal(HR[3 * n:, 3 * n:], H[n:, n:])
    np.testing.assert_array_equal(HR[:n, 3 * n:], H[:n, n:])
    assert not HR[n:3 * n, 3 * n:].any() and not HR[:n, n:3 * n].any()
