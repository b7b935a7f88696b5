"""Learned preconditioner on the GPU, through the C ABI.

* node graph and node features of D vs a direct reading of the BCOO input;
* BF16 graph-MLP forward and gradients (hand-written MFMA kernels) vs the
  NumPy oracle (oracle/learned.py) with the same bf16 rounding points;
* flexible PCG with every correction rejected (NaN weights) is bit-identical
  to the plain host-driven PCG over the same operators;
* the committed trained weights converge on held-out systems of the
  training family to rtol 1e-8 (true residual).
Tolerances are stated per test.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import _abi, learned, plugin, systems  # noqa: E402
from oracle import learned as OL  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def c1_handle():
    s = systems.cube_system(seed=0)
    h = plugin.Handle(plugin.make_config(learned.C1_SCENE, **learned.C1_EXT))
    A = plugin.BCOO3.from_system(s, DEV)
    h.setup(A, s.qo, s.qm, s.fo, s.fn)
    return s, h, A


def _fem_graph_and_blocks(s):
    """Neighbour sets and diagonal 3x3 blocks of D read from the BCOO triplets."""
    nb0 = s.fo // 3
    nb = s.fn // 3
    nbrs = [set() for _ in range(nb)]
    D = np.zeros((nb, 3, 3))
    for r, c, v in zip(s.rows, s.cols, s.vals):
        i, j = int(r) - nb0, int(c) - nb0
        if not (0 <= i < nb and 0 <= j < nb):
            continue
        blk = np.asarray(v).reshape(3, 3).T  # column-major storage
        if i == j:
            D[i] += blk
        else:
            nbrs[i].add(j)
            nbrs[j].add(i)
    return nbrs, D


def test_node_graph_and_features(c1_handle):
    s, h, _ = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    ptr, nbr = ptr.cpu().numpy(), nbr.cpu().numpy()
    nbrs, D = _fem_graph_and_blocks(s)
    assert len(ptr) == len(nbrs) + 1
    for i, want in enumerate(nbrs):  # exact
        assert sorted(nbr[ptr[i]:ptr[i + 1]].tolist()) == sorted(want), i
    dfeat, sdiag = learned.node_features(h, DEV)
    x, sref, _ = OL.node_features(D, np.ones(3 * len(nbrs)))
    np.testing.assert_allclose(sdiag.cpu().numpy(), sref, rtol=1e-6)
    np.testing.assert_allclose(dfeat.cpu().numpy(), x[:, :9], rtol=1e-5, atol=1e-6)  # fp32 storage


def _random_batch(h, B, seed):
    rng = np.random.default_rng(seed)
    info = h.setup_info()
    dfeat, _ = learned.node_features(h, DEV)
    r = torch.as_tensor(rng.normal(size=(B, 3 * info["fem_nodes"])), device=DEV)
    X, rho = learned.features(dfeat, r)
    return X, rng


def test_mlp_forward_matches_oracle(c1_handle):
    """Relative error of O (norm) <= 2e-2: both sides round the same
    activations to bf16, but FP32 (GPU) vs FP64 (oracle) sums may land on
    different sides of a bf16 rounding boundary."""
    _, h, _ = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    net = learned.GraphMLP(ptr, nbr, 4)
    theta = learned.init_params(5)
    rng = np.random.default_rng(6)
    theta = theta + 0.05 * rng.normal(size=theta.size).astype(np.float32) * (theta != 0)
    net.set_params(theta)
    np.testing.assert_array_equal(net.get_params(), theta)
    X, _ = _random_batch(h, 3, 7)
    O = net.forward(X).cpu().numpy()
    xo = X.float().cpu().numpy().astype(np.float64)
    P = learned.to_oracle_params(theta)
    Oref, _ = OL.forward(P, xo, ptr.cpu().numpy(), nbr.cpu().numpy(), bf=True)
    err = np.linalg.norm(O - Oref) / np.linalg.norm(Oref)
    assert err <= 2e-2, err
    assert np.isfinite(O).all()


def test_mlp_gradients_match_oracle(c1_handle):
    """Loss within 1e-2 relative; each parameter block's gradient within
    5e-2 relative (norm): the GPU additionally stores dO and the masked
    layer gradients in bf16 before the weight-gradient MFMA products."""
    _, h, _ = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    B = 3
    net = learned.GraphMLP(ptr, nbr, B)
    theta = learned.init_params(8)
    rng = np.random.default_rng(9)
    theta = (theta + 0.1 * rng.normal(size=theta.size).astype(np.float32) * (theta != 0)).astype(np.float32)
    # make the output layer large enough that every layer gets a gradient
    net.set_params(theta)
    X, rng = _random_batch(h, B, 10)
    nb = ptr.numel() - 1
    T = torch.as_tensor(rng.normal(size=(B, nb, 3)), device=DEV, dtype=torch.float32)
    den = torch.sum(T * T, dim=(1, 2))
    loss, g = net.gradients(X, T, den)
    P = learned.to_oracle_params(theta)
    xo = X.float().cpu().numpy().astype(np.float64)
    lref, gref, _ = OL.loss_and_grad(P, xo, T.cpu().numpy().astype(np.float64), ptr.cpu().numpy(),
                                     nbr.cpu().numpy(), bf=True)
    assert abs(loss - lref) <= 1e-2 * abs(lref), (loss, lref)
    G = learned.to_oracle_params(g)
    for k in gref:
        a, b = G[k], gref[k]
        if k == "W0":  # rows 12..15 see zero features
            a, b = a[:12], b[:12]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err <= 5e-2, (k, err)
    # padded output columns carry no gradient
    full = g[learned.MLP_OFFSETS["Wo"]:learned.MLP_OFFSETS["Wo"] + 64 * 32].reshape(64, 32)
    assert np.all(full[:, 3:] == 0)


def test_train_step_adam_and_loss(c1_handle):
    """Adam's first step moves every parameter by -lr * g / (|g| + 1e-8)
    with g the GPU gradient (padding untouched), and 60 steps on targets
    made by a teacher network lower the loss below half its start."""
    _, h, _ = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    B = 4
    net = learned.GraphMLP(ptr, nbr, B)
    X, rng = _random_batch(h, B, 12)
    teacher = learned.init_params(2)
    o = learned.MLP_OFFSETS["Wo"]
    teacher[o:o + 64 * 32] *= 10.0
    net.set_params(teacher)
    T = net.forward(X)
    den = torch.sum(T * T, dim=(1, 2))
    theta0 = learned.init_params(1)
    net.set_params(theta0)
    _, g = net.gradients(X, T, den)
    lr = 3e-3
    losses = [net.train_step(X, T, den, lr)]
    step = net.get_params() - theta0
    want = -lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(step, want, rtol=1e-3, atol=1e-9)
    losses += [net.train_step(X, T, den, lr) for _ in range(59)]
    assert losses[-1] < 0.5 * losses[0], losses[::10]


def test_energy_targets_match_oracle(c1_handle):
    """hsc_energy_targets vs the oracle's Schur operator: energies within
    1e-10 relative (FP64 throughout), T within 1e-5 relative (stored fp32)."""
    from oracle import sparse as osp
    s, h, _ = c1_handle
    ref = osp.SparseHSC.from_system(s, order=2, omega=0.6, aca_rtol=1e-6)
    _, sdiag = learned.node_features(h, DEV)
    sd = sdiag.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(21)
    B, n = 3, s.fn
    U = rng.normal(size=(B, n))
    O = rng.normal(size=(B, n)).astype(np.float32)
    rho = np.abs(rng.normal(size=B)) + 0.5
    den = np.abs(rng.normal(size=B)) + 1.0
    Ud = torch.as_tensor(U, device=DEV)
    Od = torch.as_tensor(O, device=DEV)
    T = torch.zeros((B, n), dtype=torch.float32, device=DEV)
    en = np.zeros(B)
    h._check(h.lib.hsc_energy_targets(h.h, C.c_void_p(Od.data_ptr()), C.c_void_p(Ud.data_ptr()),
                                      rho.ctypes.data_as(C.c_void_p), den.ctypes.data_as(C.c_void_p), B,
                                      C.c_void_p(T.data_ptr()), en.ctypes.data_as(C.c_void_p), plugin._stream(None)))
    for b in range(B):
        sc = rho[b] / np.repeat(sd, 3)
        w = U[b] - sc * O[b].astype(np.float64)
        Sw = ref.matvec(w)
        assert abs(en[b] - w @ Sw) <= 1e-10 * abs(w @ Sw)
        np.testing.assert_allclose(T[b].cpu().numpy(), O[b] + sc * Sw / den[b], rtol=1e-5, atol=1e-6)


def _plain_traced(h, b, x):
    st = _abi.Stats()
    code = h.lib.hsc_solve_traced(h.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), None, None, 0, 0, None,
                                  C.byref(st), plugin._stream(None))
    torch.cuda.synchronize()
    return code, st.as_dict()


def test_fallback_is_bit_identical_to_plain_pcg(c1_handle):
    """NaN weights: every correction is rejected, the flexible PCG then runs
    the plain recurrence exactly (same x bits, same iteration count), and
    the stats flag the fallback."""
    s, _, A = c1_handle
    b = torch.as_tensor(s.b, device=DEV)
    hp = plugin.Handle(plugin.make_config(learned.C1_SCENE, **learned.C1_EXT))
    hp.setup(A, s.qo, s.qm, s.fo, s.fn)
    xp = torch.zeros_like(b)
    code, stp = _plain_traced(hp, b, xp)
    assert code == 0 and stp["fallback"] == 0
    hl = plugin.Handle(plugin.make_config(learned.C1_SCENE, preconditioner="learned", **learned.C1_EXT))
    hl.setup(A, s.qo, s.qm, s.fo, s.fn)
    w = torch.full((learned.param_count(),), float("nan"), dtype=torch.float32, device=DEV)
    hl._check(hl.lib.hsc_set_learned_weights(hl.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
    xl = torch.zeros_like(b)
    stl = hl.solve(b, xl)
    torch.cuda.synchronize()
    assert stl["fallback"] == 1 and stl["preconditioner"] == 1
    assert stl["pcg_iterations"] == stp["pcg_iterations"]
    assert torch.equal(xl, xp)


def test_learned_requires_weights(c1_handle):
    s, _, A = c1_handle
    hl = plugin.Handle(plugin.make_config(learned.C1_SCENE, preconditioner="learned", **learned.C1_EXT))
    hl.setup(A, s.qo, s.qm, s.fo, s.fn)
    b = torch.as_tensor(s.b, device=DEV)
    x = torch.zeros_like(b)
    with pytest.raises(plugin.HSCError) as e:
        hl.solve(b, x)
    assert e.value.name == "HSC_ERR_INVALID_ARGUMENT"
    bad = torch.zeros(7, dtype=torch.float32, device=DEV)
    assert hl.lib.hsc_set_learned_weights(hl.h, C.c_void_p(bad.data_ptr()), 28) == _abi.HSC_ERR_INVALID_ARGUMENT


@pytest.mark.skipif(not learned.WEIGHTS.exists(), reason="no trained weights committed")
def test_trained_weights_converge():
    """Held-out systems of the training family (seeds 200-203, never
    harvested) and the 2,178-DOF system itself: flexible PCG with the
    committed network converges to rtol 1e-8 (recursive and true residual)
    and its full-system residual matches the plain solve's (both are limited
    by the same ACA rank truncation).  Iteration counts are reported, not
    asserted: the network's correction lowers the one-step energy error but
    costs Krylov conjugacy."""
    theta = learned.load_weights()
    rows = learned.evaluate(theta, [200, 201, 202, 203, "c1"], device=DEV)
    for r in rows:
        print(r["seed"], "iterations plain", r["plain"]["iterations"], "learned", r["learned"]["iterations"])
        assert r["plain"]["status"] == 0 and r["learned"]["status"] == 0, r
        assert r["learned"]["schur_rel"] <= 1e-8, r
        assert r["learned"]["full_rel"] <= 1.5 * r["plain"]["full_rel"] + 1e-12, r
