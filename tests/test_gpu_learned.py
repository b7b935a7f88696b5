"""Learned polynomial preconditioner on the GPU, through the C ABI.

* node graph of D vs a direct reading of the BCOO input, and the network's
  input features vs the oracle's statement (oracle/learned.py);
* BF16 graph-MLP forward, gradients and Adam (hand-written MFMA kernels) vs
  the NumPy oracle with the same bf16 rounding points;
* the interval the committed network picks, and the preconditioner it
  makes, vs the oracle (forward, node mean, Chebyshev weights, passes);
* a rejected network output falls back to the plain Neumann series bit for
  bit;
* what the trained network gains: on held-out systems (seeds 200-203 of the
  2,178-DOF family and of the torus family, and the 2,178-DOF system) fewer
  PCG iterations and less device time than the plain Neumann series, and
  no more iterations in total than the same passes at a fixed interval.
Tolerances are stated per test.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import _abi, learned, plugin, systems  # noqa: E402
from oracle import learned as OL, sparse as osp  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def c1_handle():
    """The 2,178-DOF system set up with the Chebyshev passes at a fixed
    interval (the network's inputs depend on D only)."""
    s = systems.cube_system(seed=0)
    h = learned.poly_handle(s, 16, 300.0, device=DEV)
    return s, h


def _fem_graph(s):
    """Neighbour sets of D read from the BCOO triplets."""
    nb0, nb = s.fo // 3, s.fn // 3
    nbrs = [set() for _ in range(nb)]
    for r, c in zip(s.rows, s.cols):
        i, j = int(r) - nb0, int(c) - nb0
        if 0 <= i < nb and 0 <= j < nb and i != j:
            nbrs[i].add(j)
            nbrs[j].add(i)
    return nbrs


def _oracle_features(s):
    D, _, _, act = osp.split_bcoo(s.rows, s.cols, s.vals, s.qo, s.qm, s.fo, s.fn)
    return OL.poly_features(D, np.unique(np.asarray(act) // 3))


def test_node_graph_and_features(c1_handle):
    """Graph exact; every feature within one bf16 rounding (2^-8 relative)
    of the oracle's FP64 value (the GPU sums in FP64 and stores bf16)."""
    s, h = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    ptr, nbr = ptr.cpu().numpy(), nbr.cpu().numpy()
    nbrs = _fem_graph(s)
    assert len(ptr) == len(nbrs) + 1
    for i, want in enumerate(nbrs):
        assert sorted(nbr[ptr[i]:ptr[i + 1]].tolist()) == sorted(want), i
    X = learned.node_features(h, DEV).float().cpu().numpy()
    xo, optr, onbr = _oracle_features(s)
    np.testing.assert_array_equal(optr, ptr)
    np.testing.assert_allclose(X, xo, rtol=2.0 ** -8, atol=1e-7)
    assert X[:, 12].sum() == 64          # the 8 x 8 rim patch is coupled
    assert np.all(X[:, 14:] == 0) and np.all(X[:, 11] > 0)


def _feature_batch(h, B, seed):
    rng = np.random.default_rng(seed)
    X = learned.node_features(h, DEV).float()
    Xb = X[None].repeat(B, 1, 1) + torch.as_tensor(0.1 * rng.normal(size=(B,) + tuple(X.shape)), device=DEV,
                                                   dtype=torch.float32)
    Xb[..., 14:] = 0.0
    return Xb.to(torch.bfloat16).contiguous(), rng


def test_mlp_forward_matches_oracle(c1_handle):
    """Relative error of O (norm) <= 2e-2: both sides round the same
    activations to bf16, but FP32 (GPU) vs FP64 (oracle) sums may land on
    different sides of a bf16 rounding boundary."""
    _, h = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    net = learned.GraphMLP(ptr, nbr, 4)
    theta = learned.init_params(5)
    rng = np.random.default_rng(6)
    theta = theta + 0.05 * rng.normal(size=theta.size).astype(np.float32) * (theta != 0)
    net.set_params(theta)
    np.testing.assert_array_equal(net.get_params(), theta)
    X, _ = _feature_batch(h, 3, 7)
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
    _, h = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    B = 3
    net = learned.GraphMLP(ptr, nbr, B)
    theta = learned.init_params(8)
    rng = np.random.default_rng(9)
    theta = (theta + 0.1 * rng.normal(size=theta.size).astype(np.float32) * (theta != 0)).astype(np.float32)
    net.set_params(theta)
    X, rng = _feature_batch(h, B, 10)
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
        if k == "W0":  # rows 14, 15 see zero features
            a, b = a[:14], b[:14]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert err <= 5e-2, (k, err)
    # padded output columns carry no gradient
    full = g[learned.MLP_OFFSETS["Wo"]:learned.MLP_OFFSETS["Wo"] + 64 * 32].reshape(64, 32)
    assert np.all(full[:, 3:] == 0)


def test_train_step_adam_and_loss(c1_handle):
    """Adam's first step moves every parameter by -lr * g / (|g| + 1e-8)
    with g the GPU gradient (padding untouched), and 60 steps on targets
    made by a teacher network lower the loss below half its start."""
    _, h = c1_handle
    ptr, nbr = learned.node_graph(h, DEV)
    B = 4
    net = learned.GraphMLP(ptr, nbr, B)
    X, rng = _feature_batch(h, B, 12)
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


needs_weights = pytest.mark.skipif(not learned.WEIGHTS.exists(), reason="no trained weights committed")


@needs_weights
@pytest.mark.parametrize("degree", [8, 12, 14, 16])
def test_learned_interval_and_preconditioner_match_oracle(degree):
    """The committed network on the 2,178-DOF system: its channel means
    within 2e-2 of the oracle's bf16 forward on the oracle's features, the
    ratio it picks (degree 14 interpolates channels 12 and 16) equal to the
    interpolated mean, and the preconditioner within 1e-12 of the oracle's
    passes with the Chebyshev weights of [b / ratio, b]."""
    s = systems.cube_system(seed=0)
    theta = learned.load_weights()
    h = learned.poly_handle(s, degree, preconditioner="learned", weights=theta, device=DEV)
    info = h.setup_info()
    assert info["learned_used"] == 1 and info["fallback_reason"] == 0
    xo, optr, onbr = _oracle_features(s)
    lg, u = OL.predict_log_ratio(learned.to_oracle_params(theta), OL.bf16(xo), optr, onbr, degree)
    np.testing.assert_allclose(info["learned_out"], u, atol=2e-2)
    t = min(max((degree - 8) / 4, 0), 2)
    c0 = int(t)
    c1 = min(c0 + 1, 2)
    got = np.log10(info["poly_ratio"]) - learned.CENTER
    want = (1 - (t - c0)) * info["learned_out"][c0] + (t - c0) * info["learned_out"][c1]
    assert got == pytest.approx(want, abs=1e-12)
    assert abs(np.log10(info["poly_ratio"]) - lg) <= 2e-2
    D, _, _, _ = osp.split_bcoo(s.rows, s.cols, s.vals, s.qo, s.qm, s.fo, s.fn)
    w, _, _ = osp.chebyshev_weights(info["lambda_max"], degree, info["poly_ratio"])
    ref = osp.SparsePolyJacobi(D, w)
    rng = np.random.default_rng(3)
    r = rng.normal(size=s.fn)
    z = torch.zeros(s.fn, dtype=torch.float64, device=DEV)
    h.apply_precond(torch.as_tensor(r, device=DEV), z)
    zr = ref(r)
    assert np.linalg.norm(z.cpu().numpy() - zr) <= 1e-12 * np.linalg.norm(zr)


def test_rejected_output_falls_back_to_neumann():
    """NaN weights: the output is rejected at setup (fallback_reason 2) and
    the solve runs the plain Neumann series with weight min(omega, 2/b), bit
    for bit what a Neumann handle with that omega returns at the same launch
    shape."""
    s = systems.cube_system(seed=0)
    nan = np.full(learned.param_count(), np.nan, dtype=np.float32)
    shape = dict(lanes_per_row=8, wg_per_cu=1, schedule=1)
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 3,
             "linear_system/hsc/max_iter": 5000}
    hl = plugin.Handle(plugin.make_config(scene, preconditioner="learned", **shape, **learned.C1_EXT))
    w = torch.as_tensor(nan, device=DEV)
    hl._check(hl.lib.hsc_set_learned_weights(hl.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
    A = plugin.BCOO3.from_system(s, DEV)
    hl.setup(A, s.qo, s.qm, s.fo, s.fn)
    info = hl.setup_info()
    assert info["fallback_reason"] == 2 and info["learned_used"] == 0 and info["poly_ratio"] == 0
    omega = min(0.7, 2.0 / (1.2 * info["lambda_max"]))
    b = torch.as_tensor(s.b, device=DEV)
    xl = torch.zeros_like(b)
    stl = hl.solve(b, xl)
    assert stl["fallback"] == 1 and stl["fallback_reason"] == 2 and stl["learned_used"] == 0
    hn = plugin.Handle(plugin.make_config(dict(scene, **{"linear_system/hsc/omega": omega}), **shape,
                                          **learned.C1_EXT))
    hn.setup(A, s.qo, s.qm, s.fo, s.fn)
    xn = torch.zeros_like(b)
    stn = hn.solve(b, xn)
    torch.cuda.synchronize()
    assert stl["pcg_iterations"] == stn["pcg_iterations"] and torch.equal(xl, xn)


def test_learned_requires_weights():
    s = systems.cube_system(seed=0)
    hl = plugin.Handle(plugin.make_config(learned.C1_SCENE, preconditioner="learned", **learned.C1_EXT))
    A = plugin.BCOO3.from_system(s, DEV)
    with pytest.raises(plugin.HSCError) as e:
        hl.setup(A, s.qo, s.qm, s.fo, s.fn)
    assert e.value.name == "HSC_ERR_INVALID_ARGUMENT"
    bad = torch.zeros(7, dtype=torch.float32, device=DEV)
    assert hl.lib.hsc_set_learned_weights(hl.h, C.c_void_p(bad.data_ptr()), 28) == _abi.HSC_ERR_INVALID_ARGUMENT


@needs_weights
def test_trained_network_gains():
    """Held-out systems (never swept for training): the learned polynomial at
    16 passes (mixed precision, rtol 1e-8) takes fewer PCG iterations and
    less device time than the plain Neumann series on every system, converges
    (Schur residual <= 1e-8) with a full residual no worse than 1e-6, and in
    total takes no more iterations than the same passes at either fixed
    interval ratio (100, the config default, or 1000)."""
    theta = learned.load_weights()
    rows = learned.evaluate(theta, learned.eval_systems(), device=DEV, reps=3)
    tot = {"learned": 0, "chebyshev_100": 0, "chebyshev_1000": 0}
    for r in rows:
        L, N = r["learned"], r["neumann"]
        print(r["system"], "iterations neumann", N["iterations"], "cheb100", r["chebyshev_100"]["iterations"],
              "cheb1000", r["chebyshev_1000"]["iterations"], "learned", L["iterations"], f"(ratio {L['ratio']:.0f})",
              f"ms neumann {N['pcg_ms']:.2f} learned {L['pcg_ms']:.2f}")
        assert L["status"] == 0 and L["learned_used"] == 1 and L["schur_rel"] <= 1e-8, r
        assert L["full_rel"] <= 1e-6, r
        assert L["iterations"] < N["iterations"] and L["pcg_ms"] < N["pcg_ms"], r
        for k in tot:
            tot[k] += r[k]["iterations"]
    print("total iterations", tot)
    assert tot["learned"] <= min(tot["chebyshev_100"], tot["chebyshev_1000"]), tot
