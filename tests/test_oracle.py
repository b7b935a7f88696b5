"""The oracle pinned against the reference's own saved results.

tests/fixtures/ holds the reference's result files (reproduction/results/ of
the reference): fixture_*.npz (matrices + direct solutions), experiments.json
(per-iteration true-residual histories, ACA histories, iteration counts,
refinement histories) and numerical_checks.json (13 invariants).  A dropped
term, wrong sign or index in the restatement changes these histories.
"""
import json

import numpy as np
import pytest

from oracle import hsc_core
from tests.conftest import FIXTURES, load_fixture

STIFF = [1e2, 1e4, 1e6]


@pytest.mark.parametrize("stiffness", STIFF)
def test_manufactured_system_reproduces_fixture(stiffness):
    f = load_fixture(stiffness)
    D, A, E, H, b, xs = hsc_core.manufactured_system(stiffness)
    for name, got in zip(["D", "A", "E", "H", "b", "x_star"], [D, A, E, H, b, xs]):
        ref = f[name]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() / scale <= 1e-14, name


@pytest.mark.parametrize("row", range(3))
def test_inner_pcg_histories_match_reference(experiments, row):
    sysrow = experiments["systems"][row]
    f = load_fixture(sysrow["stiffness"])
    D, A, E, b = f["D"], f["A"], f["E"], f["b"]
    n, m = len(D), len(A)
    for o in sysrow["orders"]:
        solver = hsc_core.HSC(D, A, E, order=o["order"], omega=0.7, aca_max_rank=min(m, n))
        assert solver.factors.rank == o["aca_rank"]
        np.testing.assert_allclose(solver.factors.residuals, o["aca_history"], rtol=1e-6, atol=1e-15)
        x, inner = solver.solve(b)
        ref = o["inner_Schur_PCG"]
        assert inner["iterations"] == ref["iterations"]
        np.testing.assert_allclose(inner["history"], ref["history"], rtol=1e-5)
        np.testing.assert_allclose(solver.affine_neumann.spectral_radius(), o["rho_affine"], rtol=1e-9)
        np.testing.assert_allclose(solver.deformable_preconditioner.spectral_radius(), o["rho_deformable"], rtol=1e-9)
        res = np.linalg.norm(f["H"] @ x - b)
        np.testing.assert_allclose(res, o["single_step_original_relative_residual"], rtol=1e-6)


@pytest.mark.parametrize("row", range(3))
def test_exact_affine_control_and_baseline(experiments, row):
    sysrow = experiments["systems"][row]
    f = load_fixture(sysrow["stiffness"])
    D, A, E, H, b = f["D"], f["A"], f["E"], f["H"], f["b"]
    exact = hsc_core.HSC(D, A, E, exact_affine=True)
    x, info = exact.solve(b)
    ref = sysrow["exact_affine_Schur_control"]
    assert info["iterations"] == ref["iterations"]
    np.testing.assert_allclose(info["history"], ref["history"], rtol=1e-5)
    assert np.linalg.norm(x - f["x_direct"]) / np.linalg.norm(f["x_direct"]) <= 1e-9
    n, m = len(D), len(A)
    bx, base = hsc_core.pcg(lambda z: H @ z, b, hsc_core.BlockInverse(H, [3] * (n // 3) + [12] * (m // 12)), rtol=1e-9)
    assert base["iterations"] == sysrow["baseline_mixed_block_jacobi_PCG"]["iterations"]
    np.testing.assert_allclose(base["history"], sysrow["baseline_mixed_block_jacobi_PCG"]["history"], rtol=1e-5)


def test_refinement_history_matches_reference(experiments):
    sysrow = experiments["systems"][2]
    f = load_fixture(sysrow["stiffness"])
    D, A, E, H, b = f["D"], f["A"], f["E"], f["H"], f["b"]
    o = sysrow["orders"][2]  # order 5
    solver = hsc_core.HSC(D, A, E, order=5, omega=0.7, aca_max_rank=36)
    _, outer = hsc_core.refinement(H, b, solver)
    ref = o["quadratic_refinement"]
    assert outer["steps"] == ref["steps"]
    assert outer["inner_iterations"] == ref["inner_iterations"]
    np.testing.assert_allclose(outer["history"], ref["history"], rtol=1e-4)


def test_aca_sweeps_match_reference(experiments):
    D, A, E, H, b, _ = hsc_core.manufactured_system()
    for row in experiments["ACA_fixed_rank_sweep"]:
        f = hsc_core.aca(E, max_rank=row["max_rank"])
        assert f.rank == row["actual_rank"]
        np.testing.assert_allclose(f.residuals[-1], row["coupling_relative_error"], rtol=1e-8, atol=1e-15)
    for row in experiments["ACA_tolerance_sweep_nonpaper_diagnostic"]:
        f = hsc_core.aca(E, rtol=row["aca_rtol"])
        assert f.rank == row["rank"]


def test_numerical_checks_reproduce():
    """The reference's 13 invariants (run.py:70-107), re-run on the oracle."""
    saved = json.loads((FIXTURES / "numerical_checks.json").read_text())["checks"]
    rng = np.random.default_rng(501)
    U, V = rng.normal(size=(24, 4)), rng.normal(size=(45, 4))
    E = U @ V.T
    f = hsc_core.aca(E, rtol=1e-12)
    got = {}
    got["ACA exact rank-4 reconstruction"] = np.linalg.norm(E - f.U @ f.V.T) / np.linalg.norm(E)
    got["ACA recovered rank"] = abs(f.rank - 4)
    got["ACA fixed-rank budget respected"] = abs(hsc_core.aca(E, max_rank=3).rank - 3)
    got["zero coupling rank"] = hsc_core.aca(np.zeros((12, 30))).rank
    A = rng.normal(size=(24, 24))
    A = A.T @ A + 50 * np.eye(24)
    op = hsc_core.Neumann(A, 12, 5, .7)
    T = np.eye(24) - .7 * op.Binv(A)
    expected = sum(np.linalg.matrix_power(T, k) for k in range(5)) @ (.7 * op.Binv(np.eye(24)))
    actual = op(np.eye(24))
    got["Neumann recurrence vs explicit polynomial"] = np.max(abs(expected - actual))
    got["Neumann I-MA=T^r identity"] = np.max(abs(np.eye(24) - actual @ A - np.linalg.matrix_power(T, 5)))
    got["Neumann inverse symmetry"] = np.max(abs(actual - actual.T))
    D, A, E, H, b, _ = hsc_core.manufactured_system(stiffness=1e4, nx=4, ny=2, nz=2)
    solver = hsc_core.HSC(D, A, E, exact_affine=True)
    x, _ = solver.solve(b, rtol=1e-12)
    direct = np.linalg.solve(H, b)
    got["exact Schur vs independent dense solve"] = np.linalg.norm(x - direct) / np.linalg.norm(direct)
    got["exact Schur full-system residual"] = np.linalg.norm(H @ x - b) / np.linalg.norm(b)
    d = rng.normal(size=len(D))
    mv = (D - E.T @ np.linalg.solve(A, E)) @ d
    got["low-rank vs explicit Schur matvec"] = np.linalg.norm(solver.matvec(d) - mv) / np.linalg.norm(mv)
    for scale in [1., 1e-12]:
        x, _ = hsc_core.pcg(lambda z: H @ z, b * scale, hsc_core.BlockInverse(H, [3] * (len(D) // 3) + [12] * 3),
                            rtol=1e-10)
        got[f"PCG RHS scale {scale:g} true relative residual"] = np.linalg.norm(H @ x - b * scale) / np.linalg.norm(b * scale)
    x, _ = hsc_core.pcg(lambda z: H @ z, np.zeros(len(H)))
    got["PCG zero RHS"] = np.linalg.norm(x)
    for c in saved:
        assert got[c["name"]] <= c["tolerance"], c["name"]


def test_closed_forms():
    """Single-block Neumann = (1-(1-w)^r) A^-1; I - P_r A = T^r; even orders
    are indefinite when lambda_max(B^-1 H) > 2/w (a planted sign error in the
    series recursion breaks the first identity)."""
    rng = np.random.default_rng(7)
    M = rng.normal(size=(12, 12))
    A = M @ M.T + 12 * np.eye(12)
    for r in [1, 2, 5]:
        P = hsc_core.Neumann(A, 12, r, 0.7)(np.eye(12))
        np.testing.assert_allclose(P, (1 - 0.3 ** r) * np.linalg.inv(A), rtol=1e-10, atol=1e-14)
    A2 = rng.normal(size=(24, 24))
    A2 = A2 @ A2.T + 24 * np.eye(24)
    nm = hsc_core.Neumann(A2, 3, 4, 0.6)
    P = nm(np.eye(24))
    T = np.eye(24) - 0.6 * nm.Binv(A2)
    np.testing.assert_allclose(np.eye(24) - P @ A2, np.linalg.matrix_power(T, 4), atol=1e-10)


def test_pcg_dropin_agrees_with_core_schedule():
    """linear_hsc.cu's stopping rule (recursive residual + one true check)
    gives the same iteration counts as core.py's true-residual schedule."""
    for s in STIFF:
        f = load_fixture(s)
        solver = hsc_core.HSC(f["D"], f["A"], f["E"], exact_affine=True)
        rhs = solver.reduced_rhs(f["b"])
        _, info = hsc_core.pcg(solver.matvec, rhs, solver.deformable_preconditioner, 1e-10, 1000)
        _, it, rel = hsc_core.pcg_dropin(solver.matvec, rhs, solver.deformable_preconditioner, 1e-10, 1000)
        assert abs(it - info["iterations"]) <= 1 and rel <= 1e-10


def test_adaptive_abd_inverse_closed_form_and_squaring():
    """The drop-in plugin's ABD inverse (linear_hsc.cu:263-297).  One body:
    B = A^-1, so P_r = (1 - (1-w)^r) A^-1 and d doublings give
    (1 - (1-w)^(r 2^d)) A^-1 with residual (1-w)^(r 2^d).  Coupled bodies:
    each doubling squares I - A P (P + P(I - AP) = 2P - PAP is symmetric, so
    the symmetrisation changes nothing), so the residual falls quadratically
    to the tolerance.  A dropped doubling term, a wrong sign or order in the
    product, or a residual normalised differently fails one of these."""
    rng = np.random.default_rng(11)
    M = rng.normal(size=(12, 12))
    A = M @ M.T + 12 * np.eye(12)
    Ainv = np.linalg.inv(A)
    for tol, d_expect in ((1e-2, 0), (1e-5, 1), (1e-9, 2)):
        P, d, res = hsc_core.neumann_adaptive(A, order=5, omega=0.7, tol=tol)
        assert d == d_expect, (tol, d)
        np.testing.assert_allclose(P, (1 - 0.3 ** (5 * 2 ** d)) * Ainv, rtol=1e-9, atol=1e-15)
        assert res == pytest.approx(0.3 ** (5 * 2 ** d), rel=1e-6)
    # not adaptive: the fixed-order polynomial equals the Neumann oracle
    P0, d0, _ = hsc_core.neumann_adaptive(A, order=3, omega=0.6, tol=1e-12, adaptive=False)
    assert d0 == 0
    np.testing.assert_allclose(P0, hsc_core.Neumann(A, 12, 3, 0.6)(np.eye(12)), rtol=1e-12, atol=1e-15)
    # coupled bodies (the reference fixtures couple bodies 0-1 and 1-2)
    f = load_fixture(1e6)
    Af = f["A"]
    eye = np.eye(len(Af))
    res_hist = []
    for dd in range(3):
        # (beyond two doublings the residual reaches the round-off floor of
        # A P, about 1e-9 here, where squaring no longer describes it)
        P, d, res = hsc_core.neumann_adaptive(Af, order=2, omega=0.7, tol=0.0, max_doublings=dd, strict=False)
        assert d == dd
        res_hist.append(res)
        if dd:
            R = eye - Af @ Pprev
            assert np.linalg.norm(eye - Af @ P - R @ R) <= 1e-6 * np.linalg.norm(R @ R) + 1e-8
        Pprev = P
    assert res_hist[2] < res_hist[1] < res_hist[0]
    P, d, res = hsc_core.neumann_adaptive(Af, order=5, omega=0.7, tol=1e-5)
    assert res <= 1e-5 and d >= 1
    np.testing.assert_allclose(P, np.linalg.inv(Af), rtol=0, atol=1e-4 * np.abs(np.linalg.inv(Af)).max())


def test_adaptive_abd_inverse_failure_modes():
    """A tolerance below round-off is missed after 48 doublings (the
    plugin's 'exceeds ... after 48 doublings' failure); a non-SPD body block
    fails the Cholesky check."""
    f = load_fixture(1e4)
    with pytest.raises(hsc_core.AbdInverseError, match="after 48 doublings"):
        hsc_core.neumann_adaptive(f["A"], tol=1e-300)
    bad = f["A"].copy()
    bad[0, 0] = -1.0
    with pytest.raises(np.linalg.LinAlgError):
        hsc_core.neumann_adaptive(bad)
