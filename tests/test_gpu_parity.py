"""Parity of the HIP path (through the C ABI) with the oracle.

* operators (D, Schur S, Neumann preconditioner) vs the oracle / SciPy;
* ACA factors vs the oracle's full-pivot ACA;
* iteration counts vs the reference's own saved histories
  (tests/fixtures/experiments.json) in the core.py schedule;
* the 2,178-DOF and 78k-tet systems vs the oracle and an FP64 direct solve;
* mixed precision, batching, determinism, linearity, error paths.
Tolerances are stated per test.
"""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, systems  # noqa: E402
from oracle import hsc_core, sparse as osp  # noqa: E402
from tests.conftest import GOLDEN, load_fixture  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _setup(sysm, scene=None, **ext):
    """A set-up handle.  Unless a test names another, the ABD inverse is the
    exact one, so that solutions compare with direct solves at FP64
    tolerances; the plugin's default (block Neumann + doubling) is compared
    with its oracle in test_gpu_dropin.py."""
    ext.setdefault("abd_inverse", "exact")
    cfg = plugin.make_config(scene or {}, **ext)
    h = plugin.Handle(cfg)
    A = plugin.BCOO3.from_system(sysm, DEV)
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    return h, A


def _solve(h, sysm, check=True):
    b = torch.as_tensor(sysm.b, device=DEV)
    x = torch.zeros_like(b)
    st = h.solve(b, x, check=check)
    torch.cuda.synchronize()
    return x.cpu().numpy(), st


@pytest.fixture(scope="module")
def c1():
    return systems.cube_system(seed=0)


@pytest.fixture(scope="module")
def c1_oracle(c1):
    return osp.SparseHSC.from_system(c1, order=2, omega=0.6, aca_rtol=1e-6)


C1_SCENE = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
            "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}


def test_operators_match_oracle(c1, c1_oracle):
    h, A = _setup(c1, C1_SCENE, aca_rtol=1e-6)
    rng = np.random.default_rng(11)
    for _ in range(2):
        v = rng.normal(size=c1.fn)
        vt = torch.as_tensor(v, device=DEV)
        y = torch.zeros_like(vt)
        h.apply_D(vt, y)
        ref = c1_oracle.D @ v
        assert np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-13
        h.apply_schur(vt, y)
        ref = c1_oracle.matvec(v)
        assert np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-12
        h.apply_precond(vt, y)
        ref = c1_oracle.precond(v)
        assert np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-12


def test_time_kernel_measures_and_leaves_the_handle_usable(c1):
    """hsc_time_kernel: a positive time per launch for each standalone kernel
    and lanes-per-row or chunks-per-wavefront; the handle's next solve still
    matches the oracle (the timed kernels overwrite only scratch vectors):
    within 1e-7 of the FP64 direct solve."""
    h, A = _setup(c1, C1_SCENE, aca_rtol=1e-6)
    for k in ("spmv64_rows", "pass64_rows", "pass32_rows"):
        for G in (4, 8, 16):
            assert h.time_kernel(k, G, repeats=3) > 0.0
    for k in ("spmv64", "pass64", "pass32"):
        for chunks in (0, 1, 3, 64):
            assert h.time_kernel(k, chunks, repeats=3) > 0.0
    x, st = _solve(h, c1)
    assert st["status"] == 0
    xd = osp.direct_solve(c1)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-7


@pytest.mark.parametrize("stiffness", [1e2, 1e4, 1e6])
def test_aca_factors_match_oracle(stiffness):
    f = load_fixture(stiffness)
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    h, _ = _setup(sysm, {"linear_system/hsc/max_rank": 36})
    F = h.factors()
    D, Ad, E, act = osp.split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    ref = hsc_core.aca(E, max_rank=36)
    assert F["rank"] == ref.rank == 18
    np.testing.assert_array_equal(F["active"], act)
    np.testing.assert_allclose(F["U"], ref.U, rtol=1e-10, atol=1e-12 * np.abs(ref.U).max())
    np.testing.assert_allclose(F["V"], ref.V, rtol=1e-10, atol=1e-12 * np.abs(ref.V).max())
    np.testing.assert_allclose(F["aca_history"], ref.residuals, rtol=1e-8, atol=1e-14)
    # exact affine inverse
    np.testing.assert_allclose(F["P"], np.linalg.inv(Ad), rtol=1e-9, atol=1e-12 * np.abs(np.linalg.inv(Ad)).max())


@pytest.mark.parametrize("row", range(3))
def test_iterations_match_reference_histories(experiments, row):
    """core.py schedule (true residual each iteration), truncated Neumann A^-1
    (exact_affine=False as in experiments.json): counts equal within +-1."""
    sysrow = experiments["systems"][row]
    f = load_fixture(sysrow["stiffness"])
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    for o in sysrow["orders"]:
        scene = {"linear_system/tol_rate": 1e-10, "linear_system/hsc/neumann_order": o["order"],
                 "linear_system/hsc/max_iter": 1000, "linear_system/hsc/max_rank": 36,
                 "linear_system/hsc/adaptive_abd_inverse": 0}
        h, _ = _setup(sysm, scene, abd_inverse="neumann", true_residual_every_iteration=1, check_acceptance=0)
        x, st = _solve(h, sysm)
        ref = o["inner_Schur_PCG"]
        assert abs(st["pcg_iterations"] - ref["iterations"]) <= 1, (o["order"], st["pcg_iterations"])
        assert st["schur_relative_residual"] <= 1e-10
        dense = hsc_core.HSC(f["D"], f["A"], f["E"], order=o["order"], omega=0.7, aca_max_rank=36)
        xo, _ = dense.solve(f["b"])
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
        np.testing.assert_allclose(np.linalg.norm(f["H"] @ x - f["b"]),
                                   o["single_step_original_relative_residual"], rtol=1e-4)


@pytest.mark.parametrize("row", range(3))
def test_exact_affine_control(experiments, row):
    sysrow = experiments["systems"][row]
    f = load_fixture(sysrow["stiffness"])
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    for every in (0, 1):
        h, _ = _setup(sysm, {"linear_system/tol_rate": 1e-10, "linear_system/hsc/max_iter": 1000},
                      true_residual_every_iteration=every)
        x, st = _solve(h, sysm)
        assert abs(st["pcg_iterations"] - sysrow["exact_affine_Schur_control"]["iterations"]) <= 1
        assert np.linalg.norm(x - f["x_direct"]) / np.linalg.norm(f["x_direct"]) <= 1e-9
        assert st["full_relative_residual"] <= 1e-9


def test_cube_system_matches_oracle_and_direct(c1, c1_oracle):
    h, _ = _setup(c1, C1_SCENE, aca_rtol=1e-6)
    x, st = _solve(h, c1)
    n = c1.fn
    dx, dq, ost = c1_oracle.solve(c1.b[:n], c1.b[n:], rtol=1e-8, maxiter=5000)
    xo = np.concatenate([dx, dq])
    assert abs(st["pcg_iterations"] - ost["iterations"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-7
    xd = osp.direct_solve(c1)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-7
    # true Schur residual recomputed on the CPU
    rhs = c1_oracle.reduced_rhs(c1.b[:n], c1.b[n:])
    assert np.linalg.norm(rhs - c1_oracle.matvec(x[:n])) / np.linalg.norm(rhs) <= 1e-8 * 1.001
    assert st["rank"] == 6 and st["active_fem_dofs"] == 192


def test_vehicle_scale_system():
    """78k-tet system at full size: converged true residual, iterations vs the
    oracle's count (tests/golden/c2_seed1.json), error vs an FP64 sparse
    direct solve, bitwise determinism and linearity in b."""
    g = json.loads((GOLDEN / "c2_seed1.json").read_text())
    s = systems.torus_system(seed=1)
    assert s.info["nb"] == g["nb"] and s.info["nnzb_D"] == g["nnzb_D"]
    np.testing.assert_allclose(np.linalg.norm(s.b), g["b_norm"], rtol=1e-12)
    h, A = _setup(s, {"linear_system/tol_rate": 1e-8, "linear_system/hsc/max_iter": 20000})
    x, st = _solve(h, s)
    assert st["rank"] == g["rank"]
    assert abs(st["pcg_iterations"] - g["iterations"]) <= 1, st["pcg_iterations"]
    xd = osp.direct_solve(s)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-8
    H = s.sparse()
    assert np.linalg.norm(H @ x - s.b) / np.linalg.norm(s.b) <= 1e-5
    x2, st2 = _solve(h, s)
    assert np.array_equal(x, x2) and st2["pcg_iterations"] == st["pcg_iterations"]
    s.b = s.b * 1e-6
    x3, st3 = _solve(h, s)
    assert abs(st3["pcg_iterations"] - st["pcg_iterations"]) <= 1
    assert np.linalg.norm(x3 - 1e-6 * x) / np.linalg.norm(1e-6 * x) <= 1e-8


@pytest.mark.parametrize("G", [4, 8, 16])
@pytest.mark.parametrize("schedule", [1, 2])
def test_launch_shapes_agree(c1, c1_oracle, G, schedule):
    """Every launch shape the tuner can pick (lanes per block row x work
    split) solves the 2,178-DOF system to the same answer: iterations within
    1 of the oracle, solution within 1e-7 of it (the shapes differ only in
    the fixed order of the partial sums)."""
    h, _ = _setup(c1, C1_SCENE, aca_rtol=1e-6, lanes_per_row=G, wg_per_cu=1, schedule=schedule)
    x, st = _solve(h, c1)
    assert st["status"] == 0
    info = h.setup_info()
    assert info["schedule"] == schedule and info["lanes_per_row"] == G
    n = c1.fn
    dx, dq, ost = c1_oracle.solve(c1.b[:n], c1.b[n:], rtol=1e-8, maxiter=5000)
    xo = np.concatenate([dx, dq])
    assert abs(st["pcg_iterations"] - ost["iterations"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-7


def test_mixed_precision_reaches_fp64_tolerance():
    s = systems.torus_system(seed=1)
    h64, _ = _setup(s, {"linear_system/tol_rate": 1e-8, "linear_system/hsc/max_iter": 20000})
    x64, st64 = _solve(h64, s)
    h32, _ = _setup(s, {"linear_system/tol_rate": 1e-8, "linear_system/hsc/max_iter": 20000},
                    precision="mixed", residual_replace_every=25)
    x32, st32 = _solve(h32, s)
    assert st32["status"] == 0 and st32["schur_relative_residual"] <= 1e-8
    xd = osp.direct_solve(s)
    err = np.linalg.norm(x32 - xd) / np.linalg.norm(xd)
    assert err <= 1e-7, err
    assert st32["pcg_iterations"] <= 2 * st64["pcg_iterations"]


@pytest.mark.parametrize("E", [1e6, 1e8, 1e10])
def test_mixed_precision_stiffness_sweep(E):
    """Mixed precision on the 2,178-DOF system with FEM E in {1e6, 1e8, 1e10}
    (Neumann order 2, omega 0.6, rtol 1e-8): converges, stays within 1e-7 of
    the FP64 direct solve, and takes at most 10% more iterations than FP64."""
    s = systems.cube_system(seed=0, E=E)
    h64, _ = _setup(s, C1_SCENE, aca_rtol=1e-6)
    x64, st64 = _solve(h64, s)
    h32, _ = _setup(s, C1_SCENE, aca_rtol=1e-6, precision="mixed", residual_replace_every=25)
    x32, st32 = _solve(h32, s)
    assert st64["status"] == 0 and st32["status"] == 0
    assert st32["schur_relative_residual"] <= 1e-8
    xd = osp.direct_solve(s)
    err32 = np.linalg.norm(x32 - xd) / np.linalg.norm(xd)
    err64 = np.linalg.norm(x64 - xd) / np.linalg.norm(xd)
    print(f"E={E:g}: iterations fp64 {st64['pcg_iterations']} mixed {st32['pcg_iterations']}, "
          f"error vs direct fp64 {err64:.2e} mixed {err32:.2e}")
    assert err32 <= 1e-7 and err64 <= 1e-7
    assert st32["pcg_iterations"] <= 1.1 * st64["pcg_iterations"] + 2


def test_batched_matches_individual(c1):
    systems_ = [c1] + systems.batch_systems(seed=2, count=3, tets=(2000, 4000), tori=(1, 2))
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 3,
             "linear_system/hsc/max_iter": 20000}
    hs, bs, xs, ref = [], [], [], []
    for s in systems_:
        h, A = _setup(s, scene)
        h._A = A
        x, st = _solve(h, s)
        ref.append((x, st))
        hs.append(h)
        bs.append(torch.as_tensor(s.b, device=DEV))
        xs.append(torch.zeros_like(bs[-1]))
    out = plugin.solve_batched(hs, bs, xs)
    torch.cuda.synchronize()
    # one launch over all systems chunks the rows differently from a single
    # solve, so every dot product sums its partials in a different (fixed)
    # order: trajectories agree to rounding, iteration counts within 2
    for (xr, sr), xb, so in zip(ref, xs, out):
        assert so["status"] == 0
        assert abs(so["pcg_iterations"] - sr["pcg_iterations"]) <= 2
        xb = xb.cpu().numpy()
        assert np.linalg.norm(xb - xr) / np.linalg.norm(xr) <= 1e-7


def test_zero_rhs_and_error_paths(c1):
    h, A = _setup(c1, C1_SCENE, aca_rtol=1e-6)
    b = torch.zeros(c1.N, dtype=torch.float64, device=DEV)
    x = torch.ones_like(b)
    st = h.solve(b, x)
    assert st["pcg_iterations"] == 0 and float(x.abs().max()) == 0.0
    # max_iter = 1: explicit non-convergence (linear_hsc.cu:394)
    h1, _ = _setup(c1, dict(C1_SCENE, **{"linear_system/hsc/max_iter": 1}), aca_rtol=1e-6)
    _, st = _solve(h1, c1, check=False)
    assert st["status_name"] == "HSC_ERR_NOT_CONVERGED" and st["pcg_iterations"] == 1
    # indefinite FEM diagonal block (linear_hsc.cu:208)
    bad = systems.cube_system(seed=0)
    d = np.flatnonzero((bad.rows == bad.cols) & (bad.rows < bad.fn // 3))[5]
    bad.vals = bad.vals.copy()
    bad.vals[d] = -bad.vals[d]
    with pytest.raises(plugin.HSCError) as e:
        _setup(bad, C1_SCENE)
    assert e.value.name == "HSC_ERR_FEM_DIAGONAL"
    # non-SPD ABD block (linear_hsc.cu:269)
    bad = systems.cube_system(seed=0)
    d = np.flatnonzero((bad.rows == bad.cols) & (bad.rows >= bad.fn // 3))[0]
    bad.vals = bad.vals.copy()
    bad.vals[d] = -10 * np.abs(bad.vals[d])
    with pytest.raises(plugin.HSCError) as e:
        _setup(bad, C1_SCENE)
    assert e.value.name == "HSC_ERR_ABD_NOT_SPD"
    # NaN in the Hessian
    bad = systems.cube_system(seed=0)
    bad.vals = bad.vals.copy()
    bad.vals[100, 4] = np.nan
    with pytest.raises(plugin.HSCError) as e:
        _setup(bad, C1_SCENE)
    assert e.value.name in ("HSC_ERR_NONFINITE_FACTORS", "HSC_ERR_FEM_DIAGONAL")
    # invalid DOF layout
    with pytest.raises(plugin.HSCError) as e:
        h2 = plugin.Handle(plugin.make_config(C1_SCENE))
        h2.setup(A, c1.qo, 13, c1.fo, c1.fn)
    assert e.value.name == "HSC_ERR_INVALID_ARGUMENT"
    # solve before setup
    h3 = plugin.Handle(plugin.make_config(C1_SCENE))
    with pytest.raises(plugin.HSCError) as e:
        h3.solve(b, x)
    assert e.value.name == "HSC_ERR_NOT_SETUP"


def test_plugin_do_solve_writes_jsonl(c1, tmp_path):
    ls = plugin.LinearHSC(workspace=tmp_path, aca_rtol=1e-6)
    with pytest.raises(plugin.SimSystemException):
        ls.do_build({"linear_system/solver": "fused_pcg"})
    ls.do_build(C1_SCENE)
    A = plugin.BCOO3.from_system(c1, DEV)
    b = torch.as_tensor(c1.b, device=DEV)
    info = plugin.SolvingInfo(x=torch.zeros_like(b), b=b, A=A, qo=c1.qo, qm=c1.qm, fo=c1.fo, fn=c1.fn,
                              frame=3, newton=1, active_mesh_contacts=17, active_ground_contacts=4)
    st = ls.do_solve(info)
    assert info.iter_count == st["pcg_iterations"] > 0
    row = json.loads((tmp_path / "hsc_stats.jsonl").read_text().splitlines()[-1])
    assert row["frame"] == 3 and row["newton"] == 1 and row["accepted"] is True and row["solver"] == "linear_hsc"
    assert row["pcg_iterations"] == st["pcg_iterations"] and row["setup_ms"] > 0
    assert (row["active_mesh_contacts"], row["active_ground_contacts"]) == (17, 4)
    # the plugin's default ABD inverse: the block Neumann polynomial, doubled
    assert row["preconditioner"] == "neumann" and row["precision"] == "fp64"
    assert row["adaptive_abd_inverse"] is True and row["neumann_doublings"] >= 1
    assert row["abd_inverse_residual"] <= 1e-5
    assert row["gbps"] > 0 and row["fp64_oracle_error"] is None


@pytest.mark.parametrize("pre", ["neumann", "chebyshev"])
@pytest.mark.parametrize("abd", ["exact", "neumann"])
def test_jsonl_names_what_ran(c1, tmp_path, pre, abd):
    """The row names the preconditioner and precision that ran, and
    adaptive_abd_inverse is true exactly when the doubling of the ABD inverse
    ran (abd_inverse = neumann with adaptive_abd_inverse = 1)."""
    ls = plugin.LinearHSC(workspace=tmp_path, aca_rtol=1e-6, preconditioner=pre, abd_inverse=abd,
                          precision="mixed")
    ls.do_build(dict(C1_SCENE, **{"linear_system/hsc/neumann_order": 5 if pre == "chebyshev" else 2}))
    A = plugin.BCOO3.from_system(c1, DEV)
    b = torch.as_tensor(c1.b, device=DEV)
    info = plugin.SolvingInfo(x=torch.zeros_like(b), b=b, A=A, qo=c1.qo, qm=c1.qm, fo=c1.fo, fn=c1.fn)
    st = ls.do_solve(info)
    row = json.loads((tmp_path / "hsc_stats.jsonl").read_text().splitlines()[-1])
    assert row["preconditioner"] == pre and row["precision"] == "mixed" and row["status"] == 0
    assert row["adaptive_abd_inverse"] is (abd == "neumann")
    assert row["neumann_doublings"] == st["neumann_doublings"] and (row["neumann_doublings"] > 0) == (abd == "neumann")


def test_aca_by_body_equals_whole_e_aca():
    """The 78k-tet system's E couples each torus rim to its own hub, so the
    setup factors it body by body (one workgroup per body).  Against the
    whole-E full-pivot ACA (forced with a relative stop of 1e-300, which it
    never reaches before the round-off guard): the same rank, U V^T equal to
    1e-13 of |E|, the same residual, and solutions within 1e-10 of each
    other."""
    s = systems.torus_system(seed=1)
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/max_iter": 20000}
    hb, _ = _setup(s, scene)
    hg, _ = _setup(s, scene, aca_rtol=1e-300)
    Fb, Fg = hb.factors(), hg.factors()
    assert Fb["rank"] == Fg["rank"] == 56
    ub, ug = Fb["U"] @ Fb["V"].T, Fg["U"] @ Fg["V"].T
    assert np.abs(ub - ug).max() <= 1e-13 * np.abs(ug).max()
    assert Fb["aca_relative_error"] == pytest.approx(Fg["aca_relative_error"], rel=1e-6, abs=1e-12)
    # both histories end at the same residual; the body path's is in body order
    assert Fb["aca_history"][-1] == pytest.approx(Fg["aca_history"][-1], rel=1e-6, abs=1e-12)
    assert np.all(np.diff(Fb["aca_history"]) <= 0)
    xb, stb = _solve(hb, s)
    xg, stg = _solve(hg, s)
    assert stb["pcg_iterations"] == pytest.approx(stg["pcg_iterations"], abs=1)
    assert np.linalg.norm(xb - xg) / np.linalg.norm(xg) <= 1e-10
