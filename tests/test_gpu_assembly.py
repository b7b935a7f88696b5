"""hsc_assemble_bcoo (the assembly before the path) on the device, through
the C ABI: bitwise equal to the oracle's sequential in-input-order sums; a
generated system re-stated as general triplets assembles back to the
plugin's input and solves to the same x bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, systems  # noqa: E402
from oracle import assembly  # noqa: E402
from tests.test_assembly import _general_form  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("fold", [True, False])
def test_matches_oracle_bitwise(fold):
    """Random triplets over 40 block rows (many duplicates, both triangles)."""
    rng = np.random.default_rng(3)
    n = 5000
    rows, cols = rng.integers(0, 40, n), rng.integers(0, 40, n)
    vals = rng.normal(size=(n, 9))
    out = plugin.assemble_bcoo(_dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float64), 40,
                               fold_lower=fold)
    r, c, v = assembly.assemble_bcoo(rows, cols, vals, fold_lower=fold)
    np.testing.assert_array_equal(out.rows.cpu().numpy(), r)
    np.testing.assert_array_equal(out.cols.cpu().numpy(), c)
    np.testing.assert_array_equal(out.values.cpu().numpy(), v)


def test_single_triplet_and_all_dropped():
    """One lower triplet folds onto (col, row) transposed; with fold_lower
    off, an input of lower triplets only assembles to nothing."""
    v = np.arange(9, dtype=np.float64)[None, :]
    out = plugin.assemble_bcoo(_dev([5], torch.int32), _dev([2], torch.int32), _dev(v, torch.float64), 6)
    assert out.rows.tolist() == [2] and out.cols.tolist() == [5]
    np.testing.assert_array_equal(out.values.cpu().numpy()[0], v[0].reshape(3, 3, order="F").T.ravel(order="F"))
    rng = np.random.default_rng(9)
    rows = rng.integers(1, 30, 300)
    cols = rows - rng.integers(1, rows + 1)          # strictly lower
    out = plugin.assemble_bcoo(_dev(rows, torch.int32), _dev(cols, torch.int32),
                               _dev(rng.normal(size=(300, 9)), torch.float64), 30, fold_lower=False)
    assert out.rows.numel() == 0


def test_empty_input():
    e = torch.empty(0, dtype=torch.int32, device=DEV)
    out = plugin.assemble_bcoo(e, e, torch.empty((0, 9), dtype=torch.float64, device=DEV), 1)
    assert out.rows.numel() == 0


@pytest.mark.parametrize("both", [True, False])
def test_general_form_assembles_and_solves_like_the_original(both):
    s = systems.cube_system(seed=0)
    R, C, V = _general_form(s, np.random.default_rng(5), both)
    A = plugin.assemble_bcoo(_dev(R, torch.int32), _dev(C, torch.int32), _dev(V, torch.float64), s.N // 3,
                             fold_lower=not both)
    order = np.lexsort((s.cols, s.rows))
    np.testing.assert_array_equal(A.rows.cpu().numpy(), s.rows[order])
    np.testing.assert_array_equal(A.values.cpu().numpy(), s.vals[order])
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
             "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
    b = torch.as_tensor(s.b, device=DEV)
    xs = []
    for mat in (A, plugin.BCOO3.from_system(s, DEV)):
        h = plugin.Handle(plugin.make_config(scene, aca_rtol=1e-6))
        h.setup(mat, s.qo, s.qm, s.fo, s.fn)
        x = torch.zeros_like(b)
        assert h.solve(b, x)["status"] == 0
        xs.append(x.cpu().numpy())
    np.testing.assert_array_equal(xs[0], xs[1])
