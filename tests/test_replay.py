"""Newton-step replay (hsc_reimpl_amd/replay.py; advance_ipc.cu:410-440 and
linear_hsc.cu:400-435 of the reference): the model line search and the
direction checks against closed forms, and the command line's size check.
No GPU: the solves are covered by tests/test_gpu_replay.py."""
import numpy as np
import pytest

from hsc_reimpl_amd import mtx, replay, systems


def test_exact_newton_step_is_accepted_at_alpha_one():
    ls = replay.model_line_search(bdx=3.0, dHd=3.0)
    assert ls == {"alpha": 1.0, "trials": 1, "accepted": True, "energy_change": -1.5}


def test_overlong_step_is_halved():
    # E(1) = -b + 2b > 0, E(1/2) = -b/2 + b/2 = 0 <= 0 (energy not increased)
    ls = replay.model_line_search(bdx=2.0, dHd=8.0)
    assert (ls["alpha"], ls["trials"], ls["accepted"], ls["energy_change"]) == (0.5, 2, True, 0.0)
    # E(a) <= 0 iff a <= 2 b / dHd = 2 / 40: a = 1/32 after 6 trials
    ls = replay.model_line_search(bdx=1.0, dHd=40.0)
    assert (ls["alpha"], ls["trials"], ls["accepted"]) == (1 / 32, 6, True)


def test_ascent_direction_is_never_accepted():
    ls = replay.model_line_search(bdx=-1.0, dHd=1.0, max_iter=5)
    assert (ls["alpha"], ls["trials"], ls["accepted"]) == (1 / 16, 5, False)
    assert ls["energy_change"] == pytest.approx(1 / 16 + 0.5 / 256)


def test_direction_checks_against_closed_forms():
    rng = np.random.default_rng(0)
    M = rng.normal(size=(30, 30))
    H = M @ M.T + 30 * np.eye(30)
    b = rng.normal(size=30)
    xs = np.linalg.solve(H, b)
    exact = -0.5 * b @ xs
    r = replay.evaluate_direction(H, b, xs, 0.1, exact)
    assert r["descent_ok"] and r["residual_ok"] and r["accepted"] and r["alpha"] == 1.0
    assert r["decrease_vs_exact"] == pytest.approx(1.0, abs=1e-12)
    assert r["full_relative_residual"] < 1e-12
    # half the Newton step: E(1) = -bx/2 + bx/8 = -3 bx / 8, i.e. 3/4 of the exact decrease
    r = replay.evaluate_direction(H, b, 0.5 * xs, 0.1, exact)
    assert r["decrease_vs_exact"] == pytest.approx(0.75, abs=1e-12)
    assert r["full_relative_residual"] == pytest.approx(0.5, abs=1e-12)
    assert not r["residual_ok"]
    # the opposite direction: not a descent direction, never accepted
    r = replay.evaluate_direction(H, b, -xs, 0.1, exact)
    assert not r["descent_ok"] and not r["accepted"]


def test_cli_rejects_mismatched_sizes(tmp_path):
    s = systems.from_dense(np.eye(6) * 2, np.eye(12) * 3, np.zeros((12, 6)), np.ones(18))
    mtx.write_bcoo(tmp_path / "A.mtx", s.rows, s.cols, s.vals, s.N // 3)
    mtx.write_vector(tmp_path / "b.mtx", s.b)
    rc = replay.main([str(tmp_path / "A.mtx"), str(tmp_path / "b.mtx"), "--qo", "6", "--qm", "12", "--fo", "0",
                      "--fn", "9"])
    assert rc == 2
