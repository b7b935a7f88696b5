"""The graph-MLP oracle is pinned by central finite differences of its own
loss (float64, no bf16 rounding), and bf16() by known bit patterns."""
import numpy as np

from oracle import learned as L


def _tiny_graph(nb=12, seed=0):
    rng = np.random.default_rng(seed)
    # ring + chords
    nbrs = [set() for _ in range(nb)]
    for i in range(nb):
        for j in ((i + 1) % nb, (i + 3) % nb):
            nbrs[i].add(j)
            nbrs[j].add(i)
    ptr, idx = [0], []
    for i in range(nb):
        idx.extend(sorted(nbrs[i]))
        ptr.append(len(idx))
    return np.array(ptr), np.array(idx)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 1e-30])
    out = L.bf16(vals)
    assert out[0] == 1.0 and out[3] == -2.5 and out[4] == 0.0
    assert out[1] == 1.0  # tie -> even
    assert out[2] == 1.0 + 2 ** -7
    assert abs(out[5] - 1e-30) / 1e-30 < 2 ** -8


def test_gradients_match_finite_differences():
    ptr, idx = _tiny_graph()
    rng = np.random.default_rng(1)
    p = L.init_params(3)
    for k in p:
        p[k] = p[k] + 0.05 * rng.normal(size=p[k].shape)
    B, nb = 2, len(ptr) - 1
    x = rng.normal(size=(B, nb, L.IN))
    x[..., 14:] = 0.0
    t = rng.normal(size=(B, nb, 3))
    loss, g, _ = L.loss_and_grad(p, x, t, ptr, idx)
    eps = 1e-6
    for k in p:
        flat = p[k].ravel()
        for pos in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[pos]
            flat[pos] = old + eps
            lp, _, _ = L.loss_and_grad(p, x, t, ptr, idx)
            flat[pos] = old - eps
            lm, _, _ = L.loss_and_grad(p, x, t, ptr, idx)
            flat[pos] = old
            fd = (lp - lm) / (2 * eps)
            an = g[k].ravel()[pos]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (k, pos, fd, an)


def test_aggregate_adjoint():
    ptr, idx = _tiny_graph()
    rng = np.random.default_rng(2)
    h = rng.normal(size=(3, len(ptr) - 1, 5))
    g = rng.normal(size=h.shape)
    lhs = np.sum(L.aggregate(h, ptr, idx) * g)
    rhs = np.sum(h * L.aggregate_T(g, ptr, idx))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_poly_features_closed_form():
    """A 3-node chain with D_ii = a_i I and D_ij = c I (i, j adjacent): every
    feature has a closed form (s_i = a_i; neighbour mean of s; the off-diagonal
    Frobenius sum over |D_ii|_F = deg c / a_i)."""
    import scipy.sparse as sp
    a = np.array([4.0, 10.0, 25.0])
    c = -1.5
    D = np.zeros((9, 9))
    for i in range(3):
        D[3 * i:3 * i + 3, 3 * i:3 * i + 3] = a[i] * np.eye(3)
    for i, j in ((0, 1), (1, 2)):
        D[3 * i:3 * i + 3, 3 * j:3 * j + 3] = c * np.eye(3)
        D[3 * j:3 * j + 3, 3 * i:3 * i + 3] = c * np.eye(3)
    x, ptr, nbr = L.poly_features(sp.csr_matrix(D), [2])
    assert ptr.tolist() == [0, 1, 3, 4] and nbr.tolist() == [1, 0, 2, 1]
    np.testing.assert_allclose(x[:, :9], np.tile(np.eye(3).ravel(), (3, 1)))
    np.testing.assert_allclose(x[:, 9], 0.5 * np.log10([4 / 10, 10 / 14.5, 25 / 10]))
    np.testing.assert_allclose(x[:, 10], 0.25 * np.array([1, 2, 1]) * abs(c) / a)
    np.testing.assert_allclose(x[:, 11], [1 / 16, 2 / 16, 1 / 16])
    assert x[:, 12].tolist() == [0, 0, 1]
    np.testing.assert_allclose(x[:, 13], np.log10(3) / 6)
    assert np.all(x[:, 14:] == 0)


def test_predict_log_ratio_channels():
    """The node mean of the outputs picks the channel of the degree (8, 12,
    16) and interpolates between them: a network whose output is its bias
    (all weights zero) returns bias + 2.5 exactly."""
    p = L.init_params(0)
    for k in p:
        p[k] = np.zeros_like(p[k])
    p["bo"] = np.array([-0.5, 0.25, 1.0])
    ptr, idx = _tiny_graph()
    x = np.random.default_rng(0).normal(size=(len(ptr) - 1, L.IN))
    for deg, want in ((4, 2.0), (8, 2.0), (10, 2.375), (12, 2.75), (14, 3.125), (16, 3.5), (30, 3.5)):
        lg, u = L.predict_log_ratio(p, x, ptr, idx, deg, bf=False)
        assert lg == want and np.allclose(u, p["bo"])
