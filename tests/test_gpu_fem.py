"""hsc_fem_hessian (the FEM Hessian generator before the path) on the
device, through the C ABI: per-tet blocks against the oracle on deformed
tets (projection on and off, 1e-10 relative per block), and generator +
assembly on the device against the synthetic generator's blocks for the
rest-state cube mesh (1e-12 relative)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, systems  # noqa: E402
from oracle import fem  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
MU, LAM = 3.448e7, 3.103e8


def _t(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("project", [True, False])
def test_blocks_match_oracle_on_deformed_tets(project):
    nodes, tets = systems.box_mesh(3, 3, 2, 0.02)
    rng = np.random.default_rng(7)
    X = nodes[tets]
    Dm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1)
    Dinv = np.linalg.inv(Dm)
    vol = np.abs(np.linalg.det(Dm)) / 6
    x = nodes + 0.004 * rng.normal(size=nodes.shape)     # deformed: indefinite tets included
    mu = MU * (1 + rng.random(len(tets)))
    lam = LAM * (1 + rng.random(len(tets)))
    rows, cols, vals = plugin.fem_hessian(_t(tets, torch.int32), _t(x, torch.float64), _t(Dinv, torch.float64),
                                          _t(mu, torch.float64), _t(lam, torch.float64), _t(vol, torch.float64),
                                          dt=0.01, project_spd=project)
    rows, cols, vals = rows.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    o = 0
    for e, tet in enumerate(tets):
        ref = fem.tet_blocks(x[tet], Dinv[e], mu[e], lam[e], vol[e], dt=0.01, project=project)
        for a, b, B in ref:
            l, r = (a, b) if tet[a] <= tet[b] else (b, a)
            assert (rows[o], cols[o]) == (tet[l], tet[r])
            Bref = B if l == a else B.T
            got = vals[o].reshape(3, 3, order="F")
            assert np.abs(got - Bref).max() <= 1e-10 * max(np.abs(Bref).max(), 1e-300), (e, a, b)
            o += 1


def test_generated_and_assembled_rest_mesh_matches_generator():
    nodes, tets = systems.box_mesh(10, 10, 5, 0.02)
    E, nu = 1e8, 0.45
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    X = nodes[tets]
    Dm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1)
    n = len(tets)
    rows, cols, vals = plugin.fem_hessian(_t(tets, torch.int32), _t(nodes, torch.float64),
                                          _t(np.linalg.inv(Dm), torch.float64), _t(np.full(n, mu), torch.float64),
                                          _t(np.full(n, lam), torch.float64),
                                          _t(np.abs(np.linalg.det(Dm)) / 6, torch.float64))
    A = plugin.assemble_bcoo(rows, cols, vals, len(nodes))
    I, J, B = systems.fem_blocks(nodes, tets, E, nu)
    up = I <= J
    np.testing.assert_array_equal(A.rows.cpu().numpy(), I[up])
    np.testing.assert_array_equal(A.cols.cpu().numpy(), J[up])
    ref = B[up].transpose(0, 2, 1).reshape(-1, 9)
    np.testing.assert_allclose(A.values.cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
