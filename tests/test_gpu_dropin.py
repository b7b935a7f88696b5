"""The drop-in plugin's own defaults and failure modes on the HIP path.

* the reference's default ABD inverse: the block Neumann polynomial of
  order r extended by doubling until ||I - A P||_F / sqrt(qm) <= tol
  (linear_hsc.cu:263-297), against the oracle's restatement;
* every acceptance failure of linear_hsc.cu:400-435 forced from the host:
  a tolerance below reach (HSC_ERR_ABD_INVERSE), an indefinite fixed-order
  ABD polynomial that makes the direction ascent (HSC_ERR_NOT_DESCENT) and a
  rank cap far below the coupling rank (HSC_ERR_FULL_RESIDUAL);
* the solve allocates nothing and tunes nothing after setup, and the tune
  table persists to its file.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import _abi, plugin, systems  # noqa: E402
from oracle import hsc_core  # noqa: E402
from tests.conftest import load_fixture  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _setup(sysm, scene=None, **ext):
    h = plugin.Handle(plugin.make_config(scene or {}, **ext))
    A = plugin.BCOO3.from_system(sysm, DEV)
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    h._A = A
    return h


def _solve(h, b, check=True):
    bt = torch.as_tensor(b, device=DEV)
    x = torch.zeros_like(bt)
    st = h.solve(bt, x, check=check)
    torch.cuda.synchronize()
    return x.cpu().numpy(), st


@pytest.mark.parametrize("stiffness", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("order", [1, 5])
def test_adaptive_abd_inverse_matches_oracle(stiffness, order):
    """abd_inverse = neumann with adaptive_abd_inverse = 1 (the reference's
    default pair): P within 1e-9 of the oracle's (relative to max |P|), the
    same number of doublings (4 at order 1, 1 at order 5 on these fixtures),
    the same residual, and the solve in the core.py schedule within 1
    iteration and 1e-8 of the oracle's HSC built on that P."""
    f = load_fixture(stiffness)
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    scene = {"linear_system/tol_rate": 1e-10, "linear_system/hsc/neumann_order": order,
             "linear_system/hsc/max_iter": 1000, "linear_system/hsc/max_rank": 36,
             "linear_system/hsc/adaptive_abd_inverse": 1, "linear_system/hsc/abd_inverse_tol": 1e-5}
    h = _setup(sysm, scene, abd_inverse="neumann", true_residual_every_iteration=1, check_acceptance=0)
    Pref, dref, rref = hsc_core.neumann_adaptive(f["A"], order=order, omega=0.7, tol=1e-5)
    F = h.factors()
    assert F["neumann_doublings"] == dref and dref >= 1
    np.testing.assert_allclose(F["P"], Pref, rtol=0, atol=1e-9 * np.abs(Pref).max())
    # (after 4 doublings both sit at the ~4e-9 round-off floor of A P, where
    # the two summation orders differ by about 1%)
    assert F["abd_inverse_residual"] == pytest.approx(rref, rel=1e-5, abs=1e-10)
    x, st = _solve(h, sysm.b)
    assert st["adaptive_abd_inverse"] == 1 and st["neumann_doublings"] == dref
    dense = hsc_core.HSC(f["D"], f["A"], f["E"], order=order, omega=0.7, aca_max_rank=36, affine_P=Pref)
    xo, ost = dense.solve(f["b"], rtol=1e-10, maxiter=1000)
    assert abs(st["pcg_iterations"] - ost["iterations"]) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8


@pytest.mark.parametrize("case", ["c1", "c2"])
def test_default_config_is_the_plugin_algorithm(case):
    """hsc_config_default as the plugin's do_build leaves it (block Neumann
    ABD inverse doubled to 1e-5, Neumann series of D, FP64, drop-in stopping
    rule), through LinearHSC.do_solve, on the 2,178-DOF system (order 2,
    omega 0.6) and the 78k-tet system (order 5, omega 0.7): the same
    doublings as the oracle, PCG iterations within 1 of the oracle's HSC on
    that P, solution within 1e-7 of it; against a sparse direct solve the
    solution carries the ABD inverse's own error (measured ~6e-6, bounded
    here by 5e-5) and is accepted."""
    import json
    import tempfile
    from oracle import sparse as osp
    if case == "c1":
        s, order, omega = systems.cube_system(seed=0), 2, 0.6
        scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
                 "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
    else:
        s, order, omega = systems.torus_system(seed=1), 5, 0.7
        scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/max_iter": 20000}
    with tempfile.TemporaryDirectory() as ws:
        ls = plugin.LinearHSC(workspace=ws)
        ls.do_build(scene)
        A = plugin.BCOO3.from_system(s, DEV)
        b = torch.as_tensor(s.b, device=DEV)
        info = plugin.SolvingInfo(x=torch.zeros_like(b), b=b, A=A, qo=s.qo, qm=s.qm, fo=s.fo, fn=s.fn)
        st = ls.do_solve(info)
        row = json.loads(open(f"{ws}/hsc_stats.jsonl").read().splitlines()[-1])
    x = info.x.cpu().numpy()
    assert row["accepted"] is True and row["adaptive_abd_inverse"] is True
    D, Ad, E, act = osp.split_bcoo(s.rows, s.cols, s.vals, s.qo, s.qm, s.fo, s.fn)
    ref = osp.SparseHSC(D, Ad, E, act, order=order, omega=omega, abd_inverse="neumann")
    assert st["neumann_doublings"] == ref.doublings >= 1
    assert st["abd_inverse_residual"] == pytest.approx(ref.abd_residual, rel=1e-4)
    n = s.fn
    dx, dq, ost = ref.solve(s.b[:n], s.b[n:], rtol=1e-8, maxiter=20000)
    xo = np.concatenate([dx, dq])
    assert abs(st["pcg_iterations"] - ost["iterations"]) <= 1, (st["pcg_iterations"], ost["iterations"])
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-7
    xd = osp.direct_solve(s)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    print(f"{case}: doublings {st['neumann_doublings']} ABD residual {st['abd_inverse_residual']:.2e} "
          f"iterations {st['pcg_iterations']} (oracle {ost['iterations']}) error vs direct {err:.2e}")
    assert err <= 5e-5


def test_abd_inverse_tolerance_below_reach():
    """A tolerance below round-off is missed after 48 doublings:
    HSC_ERR_ABD_INVERSE (linear_hsc.cu:295-297), as in the oracle."""
    f = load_fixture(1e4)
    with pytest.raises(hsc_core.AbdInverseError):
        hsc_core.neumann_adaptive(f["A"], tol=1e-300)
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    with pytest.raises(plugin.HSCError) as e:
        _setup(sysm, {"linear_system/hsc/abd_inverse_tol": 1e-300}, abd_inverse="neumann")
    assert e.value.name == "HSC_ERR_ABD_INVERSE"
    assert "48 doublings" in str(e.value)


def _indefinite_abd_system():
    """The reference fixture's FEM block with its coupling scaled by 1e-3 and
    three ABD bodies tied by one stiff rank-1 coupling (weight 100), so that
    lambda_max(B^-1 A) approaches 3 > 2 / 0.7 and the fixed order-2 ABD
    polynomial is indefinite; b sits on its most negative eigenvector.  H
    itself stays SPD (min eigenvalue ~1)."""
    f = load_fixture(1e4)
    D, E = f["D"], f["E"] * 1e-3
    rng = np.random.default_rng(5)
    A2 = np.zeros((36, 36))
    for body in range(3):
        Q, _ = np.linalg.qr(rng.normal(size=(12, 12)))
        A2[12 * body:12 * body + 12, 12 * body:12 * body + 12] = Q @ np.diag(np.geomspace(1, 10, 12)) @ Q.T
    g = rng.normal(size=36)
    A2 += 100.0 * np.outer(g, g)
    P, _, _ = hsc_core.neumann_adaptive(A2, order=2, omega=0.7, adaptive=False)
    ev, V = np.linalg.eigh(P)
    b = np.concatenate([np.zeros(len(D)), V[:, 0]])
    return systems.from_dense(D, A2, E, b), P, ev[0]


def test_non_descent_direction_is_rejected():
    sysm, P, emin = _indefinite_abd_system()
    assert emin < 0
    H = sysm.dense()
    assert np.linalg.eigvalsh(H).min() > 0.5
    scene = {"linear_system/tol_rate": 1e-10, "linear_system/hsc/neumann_order": 2,
             "linear_system/hsc/adaptive_abd_inverse": 0, "linear_system/hsc/max_iter": 1000}
    # the oracle's direction is ascent: b^T x < 0
    n = sysm.fn
    dense = hsc_core.HSC(H[:n, :n], H[n:, n:], H[n:, :n], order=2, omega=0.7, affine_P=P)
    xo, _ = dense.solve(sysm.b, rtol=1e-10)
    assert sysm.b @ xo < 0
    h = _setup(sysm, scene, abd_inverse="neumann")
    x, st = _solve(h, sysm.b, check=False)
    assert st["status_name"] == "HSC_ERR_NOT_DESCENT", st
    assert st["descent"] < 0 and st["descent"] == pytest.approx(sysm.b @ xo, rel=1e-6)
    assert "non-descent" in h.lib.hsc_last_error(h.h).decode()
    # reported only when the checks are off
    h2 = _setup(sysm, scene, abd_inverse="neumann", check_acceptance=0)
    _, st2 = _solve(h2, sysm.b)
    assert st2["status"] == 0 and st2["descent"] < 0


def test_full_residual_limit_is_enforced():
    """ACA capped at rank 2 of a rank-18 coupling leaves a full residual of
    ~3e-2 (the oracle's value): above a 1e-3 limit HSC_ERR_FULL_RESIDUAL, with
    the accepted flag down; below the default 0.1 it passes."""
    f = load_fixture(1e4)
    sysm = systems.from_dense(f["D"], f["A"], f["E"], f["b"])
    dense = hsc_core.HSC(f["D"], f["A"], f["E"], exact_affine=True, aca_max_rank=2, check_schur=False)
    xo, _ = dense.solve(f["b"], rtol=1e-10)
    full_ref = np.linalg.norm(f["H"] @ xo - f["b"]) / np.linalg.norm(f["b"])
    assert 1e-3 < full_ref < 0.1
    scene = {"linear_system/tol_rate": 1e-10, "linear_system/hsc/max_rank": 2,
             "linear_system/hsc/full_residual_limit": 1e-3}
    h = _setup(sysm, scene, abd_inverse="exact")
    x, st = _solve(h, sysm.b, check=False)
    assert st["status_name"] == "HSC_ERR_FULL_RESIDUAL" and st["accepted"] == 0
    assert st["full_relative_residual"] == pytest.approx(full_ref, rel=1e-6)
    assert "exceeds" in h.lib.hsc_last_error(h.h).decode()
    h2 = _setup(sysm, dict(scene, **{"linear_system/hsc/full_residual_limit": 0.1}), abd_inverse="exact")
    _, st2 = _solve(h2, sysm.b)
    assert st2["status"] == 0 and st2["accepted"] == 1


def test_solve_allocates_and_tunes_nothing_after_setup():
    """hsc_setup picks the launch shape (timing it when the class is new)
    and sizes its scratch; hsc_solve then makes no allocation, so the first
    solve costs what later ones do (linear_hsc.cu:38-56 sizes its workspaces
    in setup too)."""
    lib = _abi.load()
    s = systems.torus_system(seed=1, n_tori=2, n_bodies=2, youngs=(1e10, 1e8))
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 16,
             "linear_system/hsc/max_iter": 5000}
    h = _setup(s, scene, precision="mixed", preconditioner="chebyshev", poly_interval_ratio=1000.0)
    n0 = lib.hsc_alloc_count()
    times = []
    for _ in range(4):
        x, st = _solve(h, s.b)
        assert st["status"] == 0
        times.append(st["pcg_ms"])
    assert lib.hsc_alloc_count() == n0
    assert times[0] <= 1.5 * float(np.median(times[1:])) + 0.2, times
    assert st["gbps"] > 0 and st["roofline_fraction"] == pytest.approx(st["gbps"] / 8000.0)


def test_tune_table_persists_to_its_file(tmp_path):
    """hsc_tune writes the class's entry to the table file; a table loaded
    from a file decides the launch shape of a later setup of that class
    (here an entry edited to another lanes-per-row value)."""
    lib = _abi.load()
    f1 = tmp_path / "tune.txt"
    try:
        assert lib.hsc_tune_file(str(f1).encode()) == 0
        s = systems.cube_system(seed=0)
        scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 3}
        h = _setup(s, scene)
        assert lib.hsc_tune(h.h, None) == 0
        lines = [ln.split() for ln in f1.read_text().splitlines() if ln.startswith("hsc-tune-v4")]
        assert lines, f1.read_text()
        entry = lines[-1]
        info = h.setup_info()
        G, wg, sched = map(int, entry[-3:])
        assert (G, sched) == (info["lanes_per_row"], info["schedule"])
        other = 16 if G != 16 else 4
        f2 = tmp_path / "tune_edited.txt"
        f2.write_text(" ".join(entry[:-3] + [str(other), str(wg), str(sched)]) + "\n")
        assert lib.hsc_tune_file(str(f2).encode()) == 0
        h2 = _setup(s, scene)
        assert h2.setup_info()["lanes_per_row"] == other
        x, st = _solve(h2, s.b)
        assert st["status"] == 0
        assert lib.hsc_tune(h2.h, None) == 0  # restore the timed entry for later tests
    finally:
        lib.hsc_tune_file(None)
