"""The assembly before the path (oracle side): general 3x3-block triplets ->
the upper-triangle BCOO of unique blocks, and the MatrixMarket files used to
replay systems.  CPU only; exact comparisons (sums of halves are exact)."""
import numpy as np

from hsc_reimpl_amd import mtx, systems
from oracle import assembly


def _dict_assemble(rows, cols, vals, fold_lower):
    acc = {}
    for r, c, v in zip(rows, cols, vals):
        B = np.asarray(v).reshape(3, 3, order="F")
        if r > c:
            if not fold_lower:
                continue
            r, c, B = c, r, B.T
        acc[(r, c)] = acc.get((r, c), np.zeros((3, 3))) + B
    keys = sorted(acc)
    return (np.array([k[0] for k in keys]), np.array([k[1] for k in keys]),
            np.array([acc[k].ravel(order="F") for k in keys]))


def _general_form(sysm, rng, both_triangles):
    """The system's blocks re-stated as general triplets: off-diagonal blocks
    mirrored into the lower triangle (both_triangles) or placed in a random
    orientation, every block split into two exact halves, order shuffled."""
    R, C, V = [], [], []
    for r, c, v in zip(sysm.rows, sysm.cols, sysm.vals):
        B = v.reshape(3, 3, order="F")
        placements = [(r, c, B)]
        if r != c:
            if both_triangles:
                placements.append((c, r, B.T))
            elif rng.random() < 0.5:
                placements = [(c, r, B.T)]
        for (pr, pc, PB) in placements:
            half = 0.5 * PB
            for part in (half, PB - half):
                R.append(pr)
                C.append(pc)
                V.append(part.ravel(order="F"))
    perm = rng.permutation(len(R))
    return np.array(R)[perm], np.array(C)[perm], np.array(V)[perm]


def test_oracle_matches_dictionary_accumulation():
    rng = np.random.default_rng(0)
    n = 400
    rows, cols = rng.integers(0, 30, n), rng.integers(0, 30, n)
    vals = rng.normal(size=(n, 9))
    for fold in (True, False):
        r, c, v = assembly.assemble_bcoo(rows, cols, vals, fold_lower=fold)
        rd, cd, vd = _dict_assemble(rows, cols, vals, fold)
        np.testing.assert_array_equal(r, rd)
        np.testing.assert_array_equal(c, cd)
        np.testing.assert_allclose(v, vd, rtol=1e-15, atol=1e-15)
        assert np.all(r <= c)


def test_general_form_round_trips_to_the_plugin_input():
    s = systems.cube_system(seed=0)
    rng = np.random.default_rng(1)
    order = np.lexsort((s.cols, s.rows))
    for both in (True, False):
        R, C, V = _general_form(s, rng, both)
        r, c, v = assembly.assemble_bcoo(R, C, V, fold_lower=not both)
        np.testing.assert_array_equal(r, s.rows[order])
        np.testing.assert_array_equal(c, s.cols[order])
        np.testing.assert_array_equal(v, s.vals[order])


def test_matrix_market_round_trip(tmp_path):
    s = systems.cube_system(seed=0)
    mtx.write_bcoo(tmp_path / "A.mtx", s.rows, s.cols, s.vals, s.N // 3)
    mtx.write_vector(tmp_path / "b.mtx", s.b)
    r, c, v, nbr = mtx.read_bcoo(tmp_path / "A.mtx")
    assert nbr == s.N // 3
    np.testing.assert_array_equal(r, s.rows)
    np.testing.assert_array_equal(c, s.cols)
    np.testing.assert_array_equal(v, s.vals)
    np.testing.assert_array_equal(mtx.read_vector(tmp_path / "b.mtx"), s.b)
    # the dumped scalar entries are H's stored block triangle (diagonal blocks
    # whole): H = S + (S - block-diagonal part of S)^T
    data = np.loadtxt(tmp_path / "A.mtx", skiprows=2)
    S = np.zeros((s.N, s.N))
    S[data[:, 0].astype(int) - 1, data[:, 1].astype(int) - 1] = data[:, 2]
    blk = np.arange(s.N) // 3
    Sd = np.where(blk[:, None] == blk[None, :], S, 0.0)
    np.testing.assert_array_equal(S + (S - Sd).T, s.dense())
