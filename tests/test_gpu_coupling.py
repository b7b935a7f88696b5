"""hsc_abd_fem_coupling (the ABD-FEM coupling blocks before the path) on the
device, through the C ABI: triplets bitwise equal to the oracle (fixed body
and vertices, block offsets); a contact scene's coupling generated and
assembled on the device, with the rest of its blocks, equals the global
pullback G^T H_x G and solves through the plugin like the direct solve;
indices out of range and empty input."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import _abi, plugin, systems  # noqa: E402
from oracle import coupling  # noqa: E402
from tests.test_coupling_oracle import _probed_jacobian  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _t(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _device_coupling(sc, abd_off=None, fem_off=None):
    return plugin.abd_fem_coupling(_t(sc["abd_vertex"], torch.int32), _t(sc["fem_vertex"], torch.int32),
                                   _t(sc["hess"], torch.float64), _t(sc["v2b"], torch.int32),
                                   _t(sc["xbar"], torch.float64), _t(sc["body_fixed"], torch.int32),
                                   _t(sc["vertex_fixed"], torch.int32),
                                   sc["abd_block_offset"] if abd_off is None else abd_off,
                                   sc["fem_block_offset"] if fem_off is None else fem_off)


@pytest.mark.parametrize("offsets", [(None, None), (5, 1000)])
def test_matches_oracle_bitwise(offsets):
    sc = systems.contact_scene(seed=5, n_contacts=3000, fixed_body=2, n_fixed_vertices=7)
    rows, cols, vals = _device_coupling(sc, *offsets)
    ao = sc["abd_block_offset"] if offsets[0] is None else offsets[0]
    fo = sc["fem_block_offset"] if offsets[1] is None else offsets[1]
    r, c, v = coupling.coupling_blocks(sc["abd_vertex"], sc["fem_vertex"], sc["hess"], sc["v2b"], sc["xbar"],
                                       sc["body_fixed"], sc["vertex_fixed"], ao, fo)
    np.testing.assert_array_equal(rows.cpu().numpy(), r)
    np.testing.assert_array_equal(cols.cpu().numpy(), c)
    np.testing.assert_array_equal(vals.cpu().numpy(), v)


def test_contact_scene_assembles_to_the_pullback_and_solves():
    sc = systems.contact_scene(seed=4, fixed_body=1, n_fixed_vertices=5)
    rows, cols, vals = _device_coupling(sc)
    A = plugin.assemble_bcoo(torch.cat([_t(sc["rows"], torch.int32), rows]),
                             torch.cat([_t(sc["cols"], torch.int32), cols]),
                             torch.cat([_t(sc["vals"], torch.float64), vals]), sc["N"] // 3)
    s = systems.HSCSystem(A.rows.cpu().numpy(), A.cols.cpu().numpy(), A.values.cpu().numpy(), sc["b"],
                          fo=sc["fo"], fn=sc["fn"], qo=sc["qo"], qm=sc["qm"])
    H = s.dense()
    # the coupling block against the global pullback of the contacts
    nb, nv, nf = len(sc["body_fixed"]), len(sc["v2b"]), sc["fn"] // 3
    G = np.zeros((3 * nv, 12 * nb))
    for v in range(nv):
        G[3 * v:3 * v + 3, 12 * sc["v2b"][v]:12 * sc["v2b"][v] + 12] = _probed_jacobian(sc["xbar"][v])
    Hx = np.zeros((3 * nv, 3 * nf))
    for a, f, Hc in zip(sc["abd_vertex"], sc["fem_vertex"], sc["hess"]):
        if not (sc["body_fixed"][sc["v2b"][a]] or sc["vertex_fixed"][f]):
            Hx[3 * a:3 * a + 3, 3 * f:3 * f + 3] += Hc
    E = G.T @ Hx
    np.testing.assert_allclose(H[:sc["qm"], sc["fo"]:], E, rtol=0, atol=1e-12 * np.abs(E).max())
    # solve through the plugin
    x_ref = np.linalg.solve(H, sc["b"])
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
             "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
    h = plugin.Handle(plugin.make_config(scene, aca_rtol=1e-6, abd_inverse="exact"))
    h.setup(A, sc["qo"], sc["qm"], sc["fo"], sc["fn"])
    info = h.setup_info()
    assert info["rank"] <= 24          # two free bodies: rank <= 24
    b = _t(sc["b"], torch.float64)
    x = torch.zeros_like(b)
    assert h.solve(b, x)["status"] == 0
    x = x.cpu().numpy()
    assert np.linalg.norm(H @ x - sc["b"]) <= 1e-7 * np.linalg.norm(sc["b"])
    assert np.abs(x - x_ref).max() <= 1e-6 * np.abs(x_ref).max()


def test_index_out_of_range_is_an_error_and_the_device_stays_usable():
    sc = systems.contact_scene(seed=6, n_contacts=100)
    for key, bad in (("abd_vertex", len(sc["v2b"])), ("fem_vertex", sc["fn"] // 3), ("abd_vertex", -1)):
        s2 = dict(sc)
        s2[key] = sc[key].copy()
        s2[key][17] = bad
        with pytest.raises(plugin.HSCError) as ei:
            _device_coupling(s2)
        assert ei.value.code == _abi.HSC_ERR_INVALID_ARGUMENT
    s2 = dict(sc, v2b=sc["v2b"].copy())
    s2["v2b"][sc["abd_vertex"][0]] = len(sc["body_fixed"])     # body out of range
    with pytest.raises(plugin.HSCError):
        _device_coupling(s2)
    rows, _, vals = _device_coupling(sc)
    assert rows.numel() == 400 and bool(torch.isfinite(vals).all())


def test_single_contact_matches_oracle():
    sc = systems.contact_scene(seed=7, n_contacts=1)
    rows, cols, vals = _device_coupling(sc, 3, 50)
    r, c, v = coupling.coupling_blocks(sc["abd_vertex"], sc["fem_vertex"], sc["hess"], sc["v2b"], sc["xbar"],
                                       sc["body_fixed"], sc["vertex_fixed"], 3, 50)
    np.testing.assert_array_equal(rows.cpu().numpy(), r)
    np.testing.assert_array_equal(cols.cpu().numpy(), c)
    np.testing.assert_array_equal(vals.cpu().numpy(), v)


def test_empty_input():
    sc = systems.contact_scene(seed=6, n_contacts=0)
    rows, cols, vals = _device_coupling(sc)
    assert rows.numel() == 0 and vals.shape == (0, 9)
