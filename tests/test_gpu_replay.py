"""Newton-step replay on the GPU (hsc_reimpl_amd/replay.py): a generated
system dumped to MatrixMarket and replayed through the command line, and each
solver variant's direction checked against the direct solve. At rtol 1e-8
every direction is a descent direction within the full-residual limit, the
model line search takes alpha = 1, and the model decrease matches the exact
Newton step's to 1e-6. At the libuipc default tol_rate 1e-3 the inexact
directions are still accepted at alpha = 1 with >= 99% of the exact decrease."""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import mtx, replay, systems  # noqa: E402

pytestmark = pytest.mark.gpu


def test_command_line_replay_of_a_dumped_system(tmp_path, capsys):
    s = systems.cube_system(seed=0)
    mtx.write_bcoo(tmp_path / "A.mtx", s.rows, s.cols, s.vals, s.N // 3)
    mtx.write_vector(tmp_path / "b.mtx", s.b)
    rc = replay.main([str(tmp_path / "A.mtx"), str(tmp_path / "b.mtx"), "--qo", str(s.qo), "--qm", str(s.qm),
                      "--fo", str(s.fo), "--fn", str(s.fn), "--tol-rate", "1e-8", "--variants", "plain_fp64",
                      "mixed_neumann", "new"])
    assert rc == 0
    rows = [json.loads(line) for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert [r["variant"] for r in rows] == ["plain_fp64", "mixed_neumann", "new"]
    for r in rows:
        assert r["status"] == "HSC_OK", r
        assert r["descent_ok"] and r["residual_ok"] and r["solver_accepted"]
        assert r["accepted"] and r["alpha"] == 1.0 and r["trials"] == 1
        assert abs(r["decrease_vs_exact"] - 1.0) <= 1e-6, r


@pytest.mark.parametrize("variant", ["plain_fp64", "new"])
def test_inexact_directions_at_default_tolerance(variant):
    s = systems.cube_system(seed=1)
    x_exact = np.linalg.solve(s.dense(), s.b)
    rows = replay.replay(s, [variant], {"linear_system/tol_rate": 1e-3}, x_exact=x_exact)
    r = rows[0]
    assert r["status"] == "HSC_OK" and r["descent_ok"] and r["residual_ok"]
    assert r["accepted"] and r["alpha"] == 1.0
    assert 0.99 <= r["decrease_vs_exact"] <= 1.0 + 1e-9, r
