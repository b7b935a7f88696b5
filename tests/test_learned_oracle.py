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
    x[..., 12:] = 0.0
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
