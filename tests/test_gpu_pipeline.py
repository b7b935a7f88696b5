"""The caller-side producers and the solver composed on the device: the FEM
Hessian of a deformed mesh (hsc_fem_hessian), ABD-FEM contact couplings
(hsc_abd_fem_coupling), the ABD blocks and mass / contact diagonals, all
assembled by hsc_assemble_bcoo into the plugin's input, then solved by plain
FP64 and by the new solver.  The assembled matrix is checked against the same
matrix built on the CPU from the oracles (oracle.fem tet blocks,
oracle.coupling blocks, oracle.assembly; 1e-10 relative), the solutions
against a dense direct solve."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, replay, systems  # noqa: E402
from oracle import assembly, coupling, fem  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
E_MOD, NU = 1e8, 0.45


def _t(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _scene():
    sc = systems.contact_scene(seed=8, fem_stiffness=False, n_fixed_vertices=3)
    nodes, tets = sc["nodes"], sc["tets"]
    x = nodes + 0.001 * np.random.default_rng(3).normal(size=nodes.shape)
    X = nodes[tets]
    Dm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1)
    Dinv = np.linalg.inv(Dm)
    vol = np.abs(np.linalg.det(Dm)) / 6
    n = len(tets)
    mu = np.full(n, E_MOD / (2 * (1 + NU)))
    lam = np.full(n, E_MOD * NU / ((1 + NU) * (1 - 2 * NU)))
    return sc, x, Dinv, vol, mu, lam


def _cpu_matrix(sc, x, Dinv, vol, mu, lam):
    off = sc["fem_block_offset"]
    R, C, V = [sc["rows"]], [sc["cols"]], [sc["vals"]]
    fr, fc, fv = [], [], []
    for e, tet in enumerate(sc["tets"]):
        for a, b, B in fem.tet_blocks(x[tet], Dinv[e], mu[e], lam[e], vol[e], dt=1.0, project=True):
            lo = tet[a] <= tet[b]
            fr.append(off + (tet[a] if lo else tet[b]))
            fc.append(off + (tet[b] if lo else tet[a]))
            fv.append((B if lo else B.T).ravel(order="F"))
    R.append(np.array(fr))
    C.append(np.array(fc))
    V.append(np.array(fv))
    r, c, v = coupling.coupling_blocks(sc["abd_vertex"], sc["fem_vertex"], sc["hess"], sc["v2b"], sc["xbar"],
                                       sc["body_fixed"], sc["vertex_fixed"], sc["abd_block_offset"], off)
    R.append(r)
    C.append(c)
    V.append(v)
    return assembly.assemble_bcoo(np.concatenate(R), np.concatenate(C), np.concatenate(V), fold_lower=True)


def test_device_pipeline_matches_cpu_and_solves():
    sc, x, Dinv, vol, mu, lam = _scene()
    off = sc["fem_block_offset"]
    fr, fc, fv = plugin.fem_hessian(_t(sc["tets"], torch.int32), _t(x, torch.float64), _t(Dinv, torch.float64),
                                    _t(mu, torch.float64), _t(lam, torch.float64), _t(vol, torch.float64), dt=1.0)
    cr, cc, cv = plugin.abd_fem_coupling(_t(sc["abd_vertex"], torch.int32), _t(sc["fem_vertex"], torch.int32),
                                         _t(sc["hess"], torch.float64), _t(sc["v2b"], torch.int32),
                                         _t(sc["xbar"], torch.float64), _t(sc["body_fixed"], torch.int32),
                                         _t(sc["vertex_fixed"], torch.int32), sc["abd_block_offset"], off)
    A = plugin.assemble_bcoo(torch.cat([_t(sc["rows"], torch.int32), fr + off, cr]),
                             torch.cat([_t(sc["cols"], torch.int32), fc + off, cc]),
                             torch.cat([_t(sc["vals"], torch.float64), fv, cv]), sc["N"] // 3)
    r, c, v = _cpu_matrix(sc, x, Dinv, vol, mu, lam)
    np.testing.assert_array_equal(A.rows.cpu().numpy(), r)
    np.testing.assert_array_equal(A.cols.cpu().numpy(), c)
    np.testing.assert_allclose(A.values.cpu().numpy(), v, rtol=0, atol=1e-10 * np.abs(v).max())

    s = systems.HSCSystem(A.rows.cpu().numpy(), A.cols.cpu().numpy(), A.values.cpu().numpy(), sc["b"],
                          fo=sc["fo"], fn=sc["fn"], qo=sc["qo"], qm=sc["qm"])
    H = s.dense()
    x_ref = np.linalg.solve(H, sc["b"])
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
             "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
    b = _t(sc["b"], torch.float64)
    for name in ("plain_fp64", "new"):
        # exact ABD inverse: the solutions are compared with a dense solve at 1e-7
        h = replay.make_handle(scene, dict(replay.VARIANTS[name], abd_inverse="exact"), DEV, aca_rtol=1e-6)
        h.setup(A, sc["qo"], sc["qm"], sc["fo"], sc["fn"])
        xs = torch.zeros_like(b)
        st = h.solve(b, xs)
        assert st["status"] == 0, (name, st)
        xs = xs.cpu().numpy()
        assert np.linalg.norm(H @ xs - sc["b"]) <= 1e-7 * np.linalg.norm(sc["b"]), name
        assert np.abs(xs - x_ref).max() <= 1e-6 * np.abs(x_ref).max(), name
