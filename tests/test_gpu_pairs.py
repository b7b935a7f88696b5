"""Pass pairs (pass_sync = 2), through the C ABI.

Each row tile computes preconditioner pass j for its own rows and for its
halo rows (the rows of other tiles its rows reference) in LDS, then pass j+1
for its own rows from there; between pairs the tiles exchange their pass
values as tagged 64-bit words (value and round number), re-read until the
tag matches, instead of waiting on flags.  A halo row is computed with the
owner's lane split, order and helpers, so the preconditioner is the same
operator to the last bit:

* on the 78k-tet system (mixed precision, Chebyshev-weighted passes, ratio
  1000) x, the iteration count and the residuals are bitwise equal to the
  per-pass neighbour sync and to grid barriers, for 2 passes (one pair, no
  tagged exchange), 4 (two pairs, one tagged round) and 15 (pass 1 alone, then
  7 pairs: 7 tagged rounds per iteration), and the pairs actually ran;
* repeated solves are bitwise identical (every value is read with the tag
  of the round that wrote it);
* a system whose tile halo does not fit (a row of more than 20 blocks) falls
  back to one wait per pass on the device, bit for bit the per-pass result;
* the 2,178-DOF system runs pairs too, and converges to the direct solution;
* schedule 3's row tiles follow D's connected components (each torus of
  the 78k-tet system starts a tile), as a NumPy restatement of the rule
  gives them; the three pass modes stay bitwise equal on them, and the
  solve agrees with the fixed tiles' to rounding;
* with fixed tiles, a tile where two bodies meet has more than 128 halo
  rows and holds 2 lanes per halo row, so its halo takes one round (the
  bitwise tests above run on it).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, systems  # noqa: E402
from oracle import sparse as osp  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
LATENCY = {"lanes_per_row": 4, "wg_per_cu": 2, "schedule": 1}


def _setup(sysm, order, ps, ratio=1000.0, schedule=1):
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": order,
             "linear_system/hsc/max_iter": 20000}
    cfg = plugin.make_config(scene, precision="mixed", residual_replace_every=25, preconditioner="chebyshev",
                             poly_interval_ratio=ratio, abd_inverse="exact", pass_sync=ps,
                             **dict(LATENCY, schedule=schedule))
    h = plugin.Handle(cfg)
    A = plugin.BCOO3.from_system(sysm, DEV)
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    h._A = A
    return h


def _solve(h, sysm, reps=1):
    b = torch.as_tensor(sysm.b, device=DEV)
    out = []
    for _ in range(reps):
        x = torch.zeros_like(b)
        st = h.solve(b, x)
        torch.cuda.synchronize()
        out.append((x.cpu().numpy(), st))
    return out


@pytest.fixture(scope="module")
def c2():
    return systems.torus_system(seed=1)


@pytest.mark.parametrize("order,schedule", [(3, 1), (5, 1), (16, 1), (16, 3)])
def test_pairs_bitwise_equal_to_per_pass_and_grid_barriers(c2, order, schedule):
    res = {}
    for ps in (0, 1, 2):
        h = _setup(c2, order, ps, schedule=schedule)
        (x, st), = _solve(h, c2)
        info = h.setup_info()
        assert st["status"] == 0 and st["schur_relative_residual"] <= 1e-8, st
        assert info["pass_pairs"] == (1 if ps == 2 else 0), info
        assert info["pass_sync"] == (1 if ps else 0), info
        res[ps] = (x, st)
        h.close()
    for ps in (0, 2):
        assert np.array_equal(res[ps][0], res[1][0]), ps
        for key in ("pcg_iterations", "schur_relative_residual", "full_relative_residual"):
            assert res[ps][1][key] == res[1][1][key], (ps, key)


def test_pairs_repeatable(c2):
    h = _setup(c2, 16, 2)
    runs = _solve(h, c2, reps=4)
    assert h.setup_info()["pass_pairs"] == 1
    for x, st in runs[1:]:
        assert np.array_equal(x, runs[0][0])
        assert st["pcg_iterations"] == runs[0][1]["pcg_iterations"]
    h.close()


def _with_long_row(sysm, node=0, links=30, k=1e3):
    """The system plus springs (stiffness k I) from FEM node `node` to `links`
    far nodes: row `node` of D gets more than 20 blocks; H stays SPD."""
    nb = sysm.fn // 3
    far = np.linspace(nb // 2, nb - 1, links).astype(np.int64)
    rows, cols, vals = sysm.rows.copy(), sysm.cols.copy(), sysm.vals.copy()
    eye = np.eye(3).T.reshape(9)
    for c in (node, *far):
        d = np.flatnonzero((rows == c) & (cols == c))
        assert len(d) == 1
        vals[d[0]] += (links if c == node else 1) * k * eye
    rows = np.concatenate([rows, np.full(links, node, dtype=rows.dtype)])
    cols = np.concatenate([cols, far.astype(cols.dtype)])
    vals = np.concatenate([vals, np.tile(-k * eye, (links, 1))])
    order = np.lexsort((cols, rows))
    return systems.HSCSystem(rows[order], cols[order], np.ascontiguousarray(vals[order]), sysm.b.copy(),
                             fo=sysm.fo, fn=sysm.fn, qo=sysm.qo, qm=sysm.qm, info=dict(sysm.info))


def test_pairs_fall_back_when_a_halo_does_not_fit(c2):
    s = _with_long_row(c2)
    res = {}
    for ps in (1, 2):
        h = _setup(s, 16, ps)
        (x, st), = _solve(h, s)
        assert st["status"] == 0 and st["schur_relative_residual"] <= 1e-8, st
        assert h.setup_info()["pass_pairs"] == 0  # the device kept one wait per pass
        res[ps] = (x, st)
        h.close()
    assert np.array_equal(res[1][0], res[2][0])
    assert res[1][1]["pcg_iterations"] == res[2][1]["pcg_iterations"]


def test_pairs_on_the_small_system_reach_the_direct_solution():
    s = systems.cube_system(seed=0)
    h = _setup(s, 16, 2, ratio=300.0)
    (x, st), = _solve(h, s)
    assert st["status"] == 0 and h.setup_info()["pass_pairs"] == 1, (st, h.setup_info())
    xd = osp.direct_solve(s)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-7
    h.close()


def _body_tiles(sysm, rows=128):
    """The row-tile starts the latency schedule uses at 4 lanes per row: D's
    connected components by the running maximum of each row's largest
    column; a component of `rows` rows or more starts a tile, and tiles are
    cut every `rows` rows from the last such start."""
    nb = sysm.fn // 3
    r, c = sysm.rows.astype(np.int64), sysm.cols.astype(np.int64)
    fem = (r < nb) & (c < nb)
    reach = np.arange(nb)
    np.maximum.at(reach, r[fem], c[fem])
    np.maximum.at(reach, c[fem], r[fem])
    ends = np.flatnonzero(np.maximum.accumulate(reach) == np.arange(nb))
    starts = np.concatenate([[0], ends[:-1] + 1])
    big = [int(a) for a, e in zip(starts, ends) if e - a + 1 >= rows]
    segs = big if big and big[0] == 0 else [0] + big
    tiles = []
    for j, a in enumerate(segs):
        tiles += list(range(a, segs[j + 1] if j + 1 < len(segs) else nb, rows))
    return tiles


def test_row_tiles_follow_the_bodies(c2):
    """The 78k-tet system's 7 tori of 3,360 nodes start on tile boundaries:
    27 tiles each (26 of 128 rows and one of 32), 189 in all where fixed
    128-row tiles give 184; the cube (one component) keeps the fixed tiles'
    count; both match the restated rule, and the solve on them converges."""
    expect = _body_tiles(c2)
    assert len(expect) == 189 and all(a % 3360 == 0 for a in expect[::27])
    res = {}
    for sched in (1, 3):
        h = _setup(c2, 16, 2, schedule=sched)
        info = h.setup_info()
        assert info["row_tiles"] == len(expect) and info["schedule"] == sched
        (x, st), = _solve(h, c2)
        assert st["status"] == 0 and st["schur_relative_residual"] <= 1e-8 and h.setup_info()["pass_pairs"] == 1
        res[sched] = (x, st["pcg_iterations"])
        h.close()
    # other tiles, other partial sums: the same solve to rounding
    assert abs(res[1][1] - res[3][1]) <= 1
    assert np.linalg.norm(res[1][0] - res[3][0]) <= 1e-6 * np.linalg.norm(res[1][0])
    cube = systems.cube_system(seed=0)
    h = _setup(cube, 16, 2, ratio=300.0)
    assert h.setup_info()["row_tiles"] == len(_body_tiles(cube)) == (cube.fn // 3 + 127) // 128
    h.close()
