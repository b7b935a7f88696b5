"""HSC solve benchmark on one MI355X.

Workload: the 78k-tet vehicle-scale system (7 hollow tori, 78,400 Tet4,
23,520 FEM nodes = 70,560 FEM DOFs, 12 ABD bodies = 144 ABD DOFs, coupling
rank 56), Neumann order 5, omega 0.7, rtol 1e-8, synthetic data (seed 1).
A step = one full HSC solve (device setup: split, ACA, ABD inverse, coupling
factors; persistent Schur PCG; back-substitution; acceptance checks) with the
Hessian and b already resident in HBM.

Prints ONE JSON line.  `value` = solves/s of the new solver (FP32 bulk /
FP64 reductions with an FP64 true residual every 25 iterations); the plain
FP64 solver is timed in the same run and reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "HSC solves/s & PCG iters @ rtol 1e-8, 78k-tet system: new vs plain FP64"


def pcg_bytes_per_iteration(nb, nnzb, a, k, order, mixed):
    """Algorithmic HBM bytes of one Schur-PCG iteration (compulsory traffic:
    every matrix/vector element read or written once per pass).

    FP64: SpMV pass = 76*nnzb (72 B values + 4 B column) + 4*(nb+1) (row
    pointers) + 16*n (one vector gathered, one written); each of the order-1
    Neumann passes adds 72*nb (inverse diagonal blocks) + 8*n (r); stage A
    additionally reads p_old (8n) and writes p (8n); the low-rank apply reads
    W twice (2*8*a*k) plus 4*a indices; stage B reads q, p, r, x, invD and
    writes r, x, z (6*8n + 72*nb).  Mixed: 40 B per block (36 B values + 4 B
    column), 36 B per inverse block, 4 B per FP32 vector element; r, x stay FP64.
    """
    n = 3 * nb
    if not mixed:
        spmv = 76 * nnzb + 4 * (nb + 1) + 16 * n
        neu = (order - 1) * (spmv + 72 * nb + 8 * n)
        stage_a = spmv + 16 * n
        lowrank = 2 * 8 * a * k + 4 * a
        stage_b = 6 * 8 * n + 72 * nb
    else:
        spmv = 40 * nnzb + 4 * (nb + 1) + 8 * n
        neu = (order - 1) * (spmv + 36 * nb + 8 * n)
        stage_a = spmv + 8 * n
        lowrank = 2 * 4 * a * k + 4 * a
        stage_b = 4 * 8 * n + 2 * 4 * n + 36 * nb + 4 * n
    return neu + stage_a + lowrank + stage_b


def cpu_baseline(sysm, order, omega, rtol):
    """Oracle (sparse NumPy/SciPy restatement of the reference algorithm) on
    one full solve of the same system, single thread."""
    from oracle import sparse as osp
    t0 = time.perf_counter()
    ref = osp.SparseHSC.from_system(sysm, order=order, omega=omega)
    n = sysm.fn
    _, _, st = ref.solve(sysm.b[:n], sysm.b[n:], rtol=rtol, maxiter=20000)
    el = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one full solve (setup + {st['iterations']} PCG iterations) of the same 78k-tet system, "
                      f"{el:.2f} s, SciPy sparse matvecs single-threaded",
            "iterations": st["iterations"], "cpu": platform.processor() or platform.machine()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the HSC solve of one system runs on one MI355X and no "
              "multi-GPU figure is produced (see the sharding note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin, systems

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    order, omega, rtol = 5, 0.7, 1e-8
    sysm = systems.torus_system(seed=1)
    A = plugin.BCOO3.from_system(sysm, dev)
    b = torch.as_tensor(sysm.b, device=dev)
    scene = {"linear_system/tol_rate": rtol, "linear_system/hsc/neumann_order": order,
             "linear_system/hsc/omega": omega, "linear_system/hsc/max_iter": 20000}

    def run(variant):
        ext = {"precision": "mixed", "residual_replace_every": 25} if variant == "mixed" else {}
        h = plugin.Handle(plugin.make_config(scene, **ext))
        x = torch.zeros_like(b)
        for _ in range(args.warmup):
            h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
            h.solve(b, x)
        torch.cuda.synchronize()
        iters, pcg_ms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
            st = h.solve(b, x)
            iters.append(st["pcg_iterations"])
            pcg_ms.append(st["pcg_ms"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        info = h.setup_info()
        xs = x.cpu().numpy()
        return {"solves_per_s": args.steps / el, "ms_per_solve": 1e3 * el / args.steps,
                "pcg_iterations": int(np.median(iters)), "pcg_ms": float(np.median(pcg_ms)),
                "schur_relative_residual": st["schur_relative_residual"],
                "full_relative_residual": st["full_relative_residual"], "info": info, "x": xs}

    plain = run("fp64")
    new = run("mixed")
    info = plain["info"]
    nb, nnzb, a, k = info["fem_nodes"], info["nnzb"], info["active_fem_dofs"], info["rank"]
    err_vs_fp64 = float(np.linalg.norm(new["x"] - plain["x"]) / np.linalg.norm(plain["x"]))
    bpi = pcg_bytes_per_iteration(nb, nnzb, a, k, order, mixed=True)
    achieved = bpi * new["pcg_iterations"] / (new["pcg_ms"] * 1e-3) / 1e9
    bpi64 = pcg_bytes_per_iteration(nb, nnzb, a, k, order, mixed=False)
    achieved64 = bpi64 * plain["pcg_iterations"] / (plain["pcg_ms"] * 1e-3) / 1e9
    traffic = None
    prof = ROOT / "profiles" / "pcg_persistent_traffic.json"
    if prof.exists():
        traffic = json.loads(prof.read_text()).get("traffic_bytes_per_launch_mixed")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sysm, order, omega, rtol)
    line = {
        "metric": METRIC, "value": round(new["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(new["ms_per_solve"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded generator; no external data)",
        "config": {"workload": "78k-tet vehicle-scale HSC system (7 tori x 11,200 Tet4, 12 ABD bodies)",
                   "fem_nodes": nb, "fem_dofs": 3 * nb, "abd_dofs": info["abd_dofs"], "nnzb": nnzb,
                   "active_fem_dofs": a, "rank": k, "neumann_order": order, "omega": omega, "rtol": rtol,
                   "lanes_per_row": info["lanes_per_row"], "wg_per_cu": info["workgroups"]},
        "pcg_iterations": {"new": new["pcg_iterations"], "plain_fp64": plain["pcg_iterations"]},
        "plain_fp64": {"solves_per_s": round(plain["solves_per_s"], 3), "ms_per_solve": round(plain["ms_per_solve"], 3),
                       "pcg_ms": round(plain["pcg_ms"], 3), "pcg_gbs": round(achieved64, 1)},
        "speedup_vs_plain_fp64": round(new["solves_per_s"] / plain["solves_per_s"], 3),
        "fp64_solution_rel_diff": err_vs_fp64,
        "schur_relative_residual": new["schur_relative_residual"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "pcg_persistent<G, float> (whole Schur PCG, one cooperative launch)",
                     "bytes_per_iteration": bpi, "iterations": new["pcg_iterations"],
                     "kernel_ms": round(new["pcg_ms"], 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
