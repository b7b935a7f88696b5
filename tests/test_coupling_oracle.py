"""Pins of the coupling-assembly oracle (oracle/coupling.py, the checker of
hsc_abd_fem_coupling; abd_fem_linear_subsystem.cu:18-61 of the reference).

- DOF order: J^T g for a gravity vector g must have the layout of the
  reference's gravity setup (affine_body_dynamics.cu:609-613:
  G = (m g, g[0] m xbar, g[1] m xbar, g[2] m xbar)).
- J from the position map x(q) = p + A xbar probed by unit vectors of q.
- Many contacts over several bodies: the assembled coupling triplets equal
  the q-x_f block of G^T H_x G, G the global map (q, x_f) -> (x_abd, x_f)
  probed from x(q), fixed bodies / vertices masked.
"""
import numpy as np

from hsc_reimpl_amd import systems
from oracle import assembly, coupling


def _position(q, xbar):
    """The affine body map: x = p + A xbar, q = (p, rows of A)."""
    return q[:3] + q[3:].reshape(3, 3) @ xbar


def _probed_jacobian(xbar):
    base = _position(np.zeros(12), xbar)
    return np.stack([_position(np.eye(12)[j], xbar) - base for j in range(12)], axis=1)


def test_dof_order_matches_reference_gravity_layout():
    rng = np.random.default_rng(0)
    xbar, g = rng.normal(size=3), rng.normal(size=3)
    H = np.zeros((1, 3, 3))
    H[0, :, 0] = g                      # column 0 of J^T H is J^T g
    r, c, v = coupling.coupling_blocks([0], [0], H, [0], [xbar], [0], [0])
    col0 = np.concatenate([v[k].reshape(3, 3, order="F")[:, 0] for k in range(4)])
    expect = np.concatenate([g, g[0] * xbar, g[1] * xbar, g[2] * xbar])   # m = 1
    np.testing.assert_allclose(col0, expect, rtol=1e-15, atol=0)
    # a plausible slip (block k = xbar[k-1] H) breaks it
    wrong = np.concatenate([g] + [xbar[k] * g for k in range(3)])
    assert not np.allclose(col0, wrong)


def test_jacobian_matches_position_map():
    rng = np.random.default_rng(1)
    for _ in range(5):
        xbar = rng.normal(size=3)
        np.testing.assert_allclose(coupling.abd_jacobian(xbar), _probed_jacobian(xbar), rtol=0, atol=1e-15)
        H = rng.normal(size=(1, 3, 3))
        _, _, v = coupling.coupling_blocks([0], [0], H, [0], [xbar], [0], [0])
        got = np.concatenate([v[k].reshape(3, 3, order="F") for k in range(4)])
        np.testing.assert_allclose(got, _probed_jacobian(xbar).T @ H[0], rtol=1e-14, atol=1e-15)


def test_triplet_indices_and_offsets():
    sc = systems.contact_scene(seed=2, n_contacts=50, n_bodies=3, verts_per_body=10)
    r, c, _ = coupling.coupling_blocks(sc["abd_vertex"], sc["fem_vertex"], sc["hess"], sc["v2b"], sc["xbar"],
                                       sc["body_fixed"], sc["vertex_fixed"], 7, 100)
    body = sc["v2b"][sc["abd_vertex"]]
    np.testing.assert_array_equal(r.reshape(-1, 4), 7 + 4 * body[:, None] + np.arange(4)[None])
    np.testing.assert_array_equal(c.reshape(-1, 4), np.repeat(100 + sc["fem_vertex"][:, None], 4, 1))


def test_assembled_coupling_equals_global_pullback():
    sc = systems.contact_scene(seed=3, n_contacts=400, fixed_body=1, n_fixed_vertices=5)
    nb, nv, nf = len(sc["body_fixed"]), len(sc["v2b"]), sc["fn"] // 3
    # global map q -> x_abd (3 nv x 12 nb), probed from x(q) per vertex
    G = np.zeros((3 * nv, 12 * nb))
    for v in range(nv):
        b = sc["v2b"][v]
        G[3 * v:3 * v + 3, 12 * b:12 * b + 12] = _probed_jacobian(sc["xbar"][v])
    Hx = np.zeros((3 * nv, 3 * nf))      # vertex-space cross Hessian of the contacts
    for a, f, H in zip(sc["abd_vertex"], sc["fem_vertex"], sc["hess"]):
        if sc["body_fixed"][sc["v2b"][a]] or sc["vertex_fixed"][f]:
            continue
        Hx[3 * a:3 * a + 3, 3 * f:3 * f + 3] += H
    expect = G.T @ Hx
    r, c, v = coupling.coupling_blocks(sc["abd_vertex"], sc["fem_vertex"], sc["hess"], sc["v2b"], sc["xbar"],
                                       sc["body_fixed"], sc["vertex_fixed"], sc["abd_block_offset"],
                                       sc["fem_block_offset"])
    R, C, V = assembly.assemble_bcoo(r, c, v)
    got = np.zeros_like(expect)
    for rr, cc, vv in zip(R, C, V):
        got[3 * rr:3 * rr + 3, 3 * (cc - sc["fem_block_offset"]):3 * (cc - sc["fem_block_offset"]) + 3] += \
            vv.reshape(3, 3, order="F")
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    # the fixed body's rows are exactly zero
    assert not np.any(got[12:24])


def test_fixed_flags_zero_the_blocks():
    rng = np.random.default_rng(4)
    H = rng.normal(size=(3, 3, 3))
    _, _, v = coupling.coupling_blocks([0, 1, 2], [0, 1, 0], H, [0, 1, 0], rng.normal(size=(3, 3)), [0, 1],
                                       [0, 0])
    assert np.any(v[0:4]) and not np.any(v[4:8]) and np.any(v[8:12])
    _, _, v = coupling.coupling_blocks([0, 1, 2], [0, 1, 0], H, [0, 1, 0], rng.normal(size=(3, 3)), [0, 0],
                                       [1, 0])
    assert not np.any(v[0:4]) and np.any(v[4:8]) and not np.any(v[8:12])
