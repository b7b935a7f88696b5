"""Batched solves captured in a hipGraph (hsc_batch_*), through the C ABI.

* every replay of the captured batch reproduces the eager batched solve bit
  for bit, with eager work and synchronisations between replays;
* an uneven batch from the 64-system workload (seed 2) matches the same
  systems solved one at a time.
Tolerances are stated per test.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import learned, plugin, systems  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _handles(syss, scene, **ext):
    ext.setdefault("abd_inverse", "exact")  # solutions are compared with direct solves
    hs, As = [], []
    for s in syss:
        h = plugin.Handle(plugin.make_config(scene, **ext))
        A = plugin.BCOO3.from_system(s, DEV)
        h.setup(A, s.qo, s.qm, s.fo, s.fn)
        hs.append(h)
        As.append(A)
    return hs, As


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_batched(precision):
    """x after each of 3 replays is bitwise equal to the eager batched solve
    (same kernels, same launch tables, fixed-order reductions)."""
    syss = [systems.cube_system(seed=0), learned.family_system(200), learned.family_system(201)]
    hs, As = _handles(syss, learned.C1_SCENE, precision=precision, **learned.C1_EXT)
    bs = [torch.as_tensor(s.b, device=DEV) for s in syss]
    xe = [torch.zeros_like(b) for b in bs]
    ste = plugin.solve_batched(hs, bs, xe)
    torch.cuda.synchronize()
    assert all(d["status"] == 0 for d in ste)
    xg = [torch.zeros_like(b) for b in bs]
    batch = plugin.Batch(hs, bs, xg)
    for rep in range(3):
        for x in xg:
            x.fill_(float("nan") if rep % 2 else 0.0)  # eager work between replays
        torch.cuda.synchronize()
        batch.launch()
        stg = batch.stats()
        for i in range(len(syss)):
            assert torch.equal(xg[i], xe[i]), (rep, i)
            assert stg[i]["pcg_iterations"] == ste[i]["pcg_iterations"]
            assert stg[i]["schur_relative_residual"] == ste[i]["schur_relative_residual"]
            assert stg[i]["full_relative_residual"] == ste[i]["full_relative_residual"]
        # an eager single solve on another handle between replays
        h2, _ = _handles([syss[1]], learned.C1_SCENE, **learned.C1_EXT)
        x2 = torch.zeros_like(bs[1])
        assert h2[0].solve(bs[1], x2)["status"] == 0
        torch.cuda.synchronize()
    info = batch.info()
    assert info["workgroups"] > 0 and info["lanes_per_row"] in (4, 8, 16)
    batch.close()


def test_uneven_batch_matches_individual():
    """First 4 systems of the 64-system uneven workload (seed 2), mixed
    precision, order 5, omega 0.7, rtol 1e-8: every system converges in the
    captured batch; iterations within 2 of the individual solve and
    solutions within 1e-6 relative of it (both stop at 1e-8 relative Schur
    residual, with different reduction orders)."""
    syss = systems.batch_systems(seed=2, count=4)
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 5,
             "linear_system/hsc/omega": 0.7, "linear_system/hsc/max_iter": 20000}
    hs, As = _handles(syss, scene, precision="mixed", residual_replace_every=25)
    bs = [torch.as_tensor(s.b, device=DEV) for s in syss]
    xi, sti = [], []
    for h, b in zip(hs, bs):
        x = torch.zeros_like(b)
        sti.append(h.solve(b, x))
        xi.append(x)
    xg = [torch.zeros_like(b) for b in bs]
    batch = plugin.Batch(hs, bs, xg)
    batch.launch()
    stg = batch.stats()
    for i in range(len(syss)):
        assert stg[i]["status"] == 0, stg[i]
        assert stg[i]["schur_relative_residual"] <= 1e-8
        assert abs(stg[i]["pcg_iterations"] - sti[i]["pcg_iterations"]) <= 2, (i, stg[i], sti[i])
        err = (torch.linalg.vector_norm(xg[i] - xi[i]) / torch.linalg.vector_norm(xi[i])).item()
        assert err <= 1e-6, (i, err)
    batch.close()


def test_stale_batch_is_rejected():
    """A captured batch holds its handles' buffers, ranks and tile ranges:
    after one handle is set up again (or reconfigured) the replay and its
    stats return HSC_ERR_STALE_BATCH before any device work."""
    syss = [systems.cube_system(seed=0), learned.family_system(200)]
    hs, As = _handles(syss, learned.C1_SCENE, **learned.C1_EXT)
    bs = [torch.as_tensor(s.b, device=DEV) for s in syss]
    xg = [torch.zeros_like(b) for b in bs]
    batch = plugin.Batch(hs, bs, xg)
    batch.launch()
    assert all(d["status"] == 0 for d in batch.stats())
    hs[1].setup(As[1], syss[1].qo, syss[1].qm, syss[1].fo, syss[1].fn)
    with pytest.raises(plugin.HSCError) as e:
        batch.launch()
    assert e.value.name == "HSC_ERR_STALE_BATCH"
    with pytest.raises(plugin.HSCError) as e:
        batch.stats()
    assert e.value.name == "HSC_ERR_STALE_BATCH"
    batch.close()
    batch2 = plugin.Batch(hs, bs, xg)
    hs[0].set_config(hs[0].config)
    with pytest.raises(plugin.HSCError) as e:
        batch2.launch()
    assert e.value.name == "HSC_ERR_STALE_BATCH"
    batch2.close()


def test_full_uneven_batch_matches_direct_solves():
    """All 64 systems of the uneven workload (seed 2: 1-16 bodies, 1-8 tori
    of 2k-15k tets, ranks 4-16; 1.91 M DOF) in one captured graph with the
    benchmarked solver (mixed precision, 16 Chebyshev-weighted passes, ratio
    1000, rtol 1e-8): every system converges and its x is within 1e-7 of its
    own FP64 sparse direct solve (SuperLU on the full H, computed in spawned
    CPU processes)."""
    import os
    import time

    from tests import batch_reference
    t0 = time.time()
    specs = systems.batch_specs(seed=2, count=64)
    built = systems.map_processes(batch_reference.build_and_solve, specs, min(16, os.cpu_count() or 1))
    print(f"built 64 systems and their direct solves in {time.time() - t0:.1f} s", flush=True)
    syss = [s for s, _ in built]
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 16,
             "linear_system/hsc/max_iter": 5000}
    hs, As = _handles(syss, scene, precision="mixed", residual_replace_every=25, preconditioner="chebyshev",
                      poly_interval_ratio=1000.0)
    bs = [torch.as_tensor(s.b, device=DEV) for s in syss]
    xg = [torch.zeros_like(b) for b in bs]
    batch = plugin.Batch(hs, bs, xg)
    batch.launch()
    stg = batch.stats()
    print(f"batch solved at {time.time() - t0:.1f} s; iterations {min(d['pcg_iterations'] for d in stg)}-"
          f"{max(d['pcg_iterations'] for d in stg)}", flush=True)
    worst = 0.0
    for i, ((s, xd), st, x) in enumerate(zip(built, stg, xg)):
        assert st["status"] == 0, (i, st)
        assert st["schur_relative_residual"] <= 1e-8, (i, st)
        err = np.linalg.norm(x.cpu().numpy() - xd) / np.linalg.norm(xd)
        worst = max(worst, err)
        assert err <= 1e-7, (i, s.N, err)
    print(f"worst relative error vs direct {worst:.2e}", flush=True)
    batch.close()
