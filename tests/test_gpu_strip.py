"""The standalone operator kernels (hsc_apply_D, hsc_apply_pass, and through
them hsc_apply_schur / hsc_apply_precond and the solve's acceptance D x),
through the C ABI, in both mappings (config operator_mapping: 1 = row
mapping, 2 = streaming, 0 = by working set).

Streaming: a wavefront owns the rows that start in its run of 64-block
chunks and walks their blocks chunk by chunk; row sums come out of a
head-segmented scan across the wave, a row that continues past a chunk
carries its running sum into the next one, and the rows' epilogue runs from
an LDS ring once 64 rows are finished.  The cases are chosen to reach every
branch of that walk:

* the 2,178-DOF cube, the 78k-tet torus and 8 block-diagonal copies of the
  torus (several chunks per wavefront, workgroups remapped per XCD);
* a row of more than 150 blocks (one row across three chunks and more: the
  carry chain, and wavefronts whose chunks hold no row start);
* a five-node system (one partial chunk, rows shorter than a chunk);
* 16 copies of the torus with the default mapping, whose D x reads more
  than 256 MiB and so streams (the 8 copies stay on the row mapping);
* FP64 to 1e-13 of the oracle's D (scipy), the FP32 pass to FP32 rounding,
  the same result bit for bit on a repeat.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import plugin, systems  # noqa: E402
from oracle import sparse as osp  # noqa: E402
from tests.test_gpu_pairs import _with_long_row  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _tiny_system(seed=0):
    """5 FEM nodes in a chain (13 blocks: one partial chunk) and one ABD body."""
    rng = np.random.default_rng(seed)
    n, m = 15, 12
    D = np.zeros((n, n))
    for i in range(5):
        B = rng.normal(size=(3, 3))
        D[3 * i:3 * i + 3, 3 * i:3 * i + 3] = B @ B.T + 6 * np.eye(3)
        if i + 1 < 5:
            C = 0.3 * rng.normal(size=(3, 3))
            D[3 * i:3 * i + 3, 3 * i + 3:3 * i + 6] = C
            D[3 * i + 3:3 * i + 6, 3 * i:3 * i + 3] = C.T
    G = rng.normal(size=(m, m))
    A = G @ G.T + m * np.eye(m)
    E = 0.1 * rng.normal(size=(m, n))
    return systems.from_dense(D, A, E, rng.normal(size=n + m))


_CASES = {}


def _case(name):
    """Systems built once per process, on first use."""
    if name not in _CASES:
        c2 = systems.torus_system(seed=1)
        make = {"cube": lambda: systems.cube_system(seed=0), "torus": lambda: c2,
                "torus_x8": lambda: systems.replicate_fem(c2, 8), "torus_x16": lambda: systems.replicate_fem(c2, 16),
                "long_row": lambda: _with_long_row(c2, links=150), "tiny": _tiny_system}
        _CASES[name] = make[name]()
    return _CASES[name]


def _handle(sysm, mapping=0):
    cfg = plugin.make_config({"linear_system/tol_rate": 1e-8}, precision="mixed", operator_mapping=mapping)
    h = plugin.Handle(cfg)
    A = plugin.BCOO3.from_system(sysm, DEV)
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    h._A = A
    return h


def _operators(sysm):
    """The oracle's D (scipy) and its inverse diagonal blocks."""
    D, _, _, _ = osp.split_bcoo(sysm.rows, sysm.cols, sysm.vals, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    nb, f0 = sysm.fn // 3, sysm.fo // 3
    rows = np.asarray(sysm.rows, dtype=np.int64)
    d = (rows == np.asarray(sysm.cols)) & (rows >= f0) & (rows < f0 + nb)
    blocks = np.zeros((nb, 3, 3))
    blocks[rows[d] - f0] = np.asarray(sysm.vals)[d].reshape(-1, 3, 3).transpose(0, 2, 1)  # col-major storage
    return D, np.linalg.inv(blocks)


@pytest.mark.parametrize("name,mapping", [(n, m) for n in ("cube", "torus", "long_row", "tiny") for m in (1, 2)]
                         + [("torus_x8", 2), ("torus_x8", 0), ("torus_x16", 0)])
def test_operator_kernels_match_oracle(name, mapping):
    sysm = _case(name)
    h = _handle(sysm, mapping)
    streams = {1: 0, 2: 1, 0: int(name == "torus_x16")}[mapping]
    assert h.setup_info()["streaming_operators"] == streams
    D, Dinv = _operators(sysm)
    nb = sysm.fn // 3
    if name == "long_row":
        assert np.diff(D.tocsr()[:3].indptr).max() > 3 * 150  # row 0: > 150 blocks
    rng = np.random.default_rng(7)
    x, r = rng.normal(size=sysm.fn), rng.normal(size=sysm.fn)
    xt, rt = torch.as_tensor(x, device=DEV), torch.as_tensor(r, device=DEV)
    # y = D x, twice: bitwise the same
    y = torch.full_like(xt, np.nan)
    h.apply_D(xt, y)
    ref = D @ x
    y1 = y.cpu().numpy()
    assert np.linalg.norm(y1 - ref) / np.linalg.norm(ref) <= 1e-13
    h.apply_D(xt, y)
    assert np.array_equal(y.cpu().numpy(), y1)
    # one pass, FP64 and FP32
    w = 0.6
    pref = x + w * np.einsum("nij,nj->ni", Dinv, (r - ref).reshape(nb, 3)).reshape(-1)
    z = torch.full_like(xt, np.nan)
    h.apply_pass(xt, rt, z, w)
    assert np.linalg.norm(z.cpu().numpy() - pref) / np.linalg.norm(pref) <= 1e-13
    x32 = xt.float()
    z32 = torch.full_like(x32, np.nan)
    h.apply_pass(x32, rt, z32, w)
    assert np.linalg.norm(z32.double().cpu().numpy() - pref) / np.linalg.norm(pref) <= 2e-6
    h.close()


def test_apply_pass_rejects_mixed_types():
    h = _handle(_case("tiny"))
    x = torch.zeros(15, dtype=torch.float64, device=DEV)
    with pytest.raises(TypeError):
        h.apply_pass(x, x, x.float(), 0.5)
    h.close()
