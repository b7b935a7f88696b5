"""HSC solve benchmark on one MI355X.

Workload: the 78k-tet vehicle-scale system (7 hollow tori, 78,400 Tet4,
23,520 FEM nodes = 70,560 FEM DOFs, 12 ABD bodies = 144 ABD DOFs, coupling
rank 56), rtol 1e-8, synthetic data (seed 1).
A step = one full HSC solve (device setup: split, ACA, ABD inverse, coupling
factors; persistent Schur PCG; back-substitution; acceptance checks) with the
Hessian and b already resident in HBM.

Prints ONE JSON line.  `value` = solves/s of the new solver: FP32 bulk /
FP64 reductions with an FP64 true residual every 25 iterations, and the
learned preconditioner -- 16 block-Jacobi passes weighted by the Chebyshev
roots (Leja order) of the interval [b/ratio, b] whose ratio the committed
graph MLP picks at setup from D's node features, b from a power-iteration
spectrum bound.  The plain FP64 solver (the reference algorithm: Neumann
series of order 5, omega 0.7, FP64) is timed in the same run and reported
beside it, with the same passes at the hand-picked fixed ratio 1000, the
mixed-precision Neumann and the FP64 Chebyshev variants.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
NEW_PASSES, NEW_RATIO = 16, 1000.0  # the new solver's preconditioner: passes, Chebyshev interval b/a
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


def beyond_mall(sysm, dev, copies=32):
    """The standalone SpMV / block-Jacobi pass kernels with the working set
    beyond the 256 MB Infinity Cache: `copies` block-diagonal copies of the
    78k-tet system's FEM part (systems.replicate_fem), each kernel timed over
    40 back-to-back launches with HIP events (hsc_time_kernel): the streaming
    kernels the library's operators run (chunks per wavefront: the library's
    choice), and beside them the row mapping of the persistent solve at its
    best lanes-per-row; with a measured 1 GiB device-to-device copy for
    scale."""
    import torch
    from hsc_reimpl_amd import plugin, systems
    big = systems.replicate_fem(sysm, copies)
    A = plugin.BCOO3.from_system(big, dev)
    h = plugin.Handle(plugin.make_config({"linear_system/tol_rate": 1e-8}, precision="mixed"))
    h.setup(A, big.qo, big.qm, big.fo, big.fn)
    i = h.setup_info()
    nb, nnzb = i["fem_nodes"], i["nnzb"]
    n = 3 * nb
    # algorithmic bytes: values + column indices + row pointers, x gathered
    # once, y written (+ r read and the inverse diagonal blocks for a pass)
    kernels = {"spmv64": ("y = D x, FP64", 76 * nnzb + 4 * (nb + 1) + 16 * n),
               "pass64": ("y = x + w Dinv (r - D x), FP64", 76 * nnzb + 4 * (nb + 1) + 24 * n + 72 * nb),
               "pass32": ("the same pass, FP32 values/vectors, FP64 r", 40 * nnzb + 4 * (nb + 1) + 16 * n + 36 * nb)}
    out = {"system": f"{copies} copies of the 78k-tet system's FEM part", "fem_nodes": nb, "nnzb": nnzb,
           "working_set_MB": {k: round(v[1] / 1e6, 1) for k, v in kernels.items()}, "kernels": {}}
    out["row_mapped"] = {}
    for k, (what, nbytes) in kernels.items():
        h.time_kernel(k, 0, repeats=5)
        ms = h.time_kernel(k, 0, repeats=40)
        gbs = nbytes / (ms * 1e-3) / 1e9
        out["kernels"][k] = {"what": what, "mapping": "streaming: 64-block chunks per wavefront, segmented scan",
                             "us": round(1e3 * ms, 1), "bytes": nbytes, "GB/s": round(gbs, 1),
                             "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 3)}
        ms, G = min((h.time_kernel(k + "_rows", G, repeats=40), G) for G in (4, 8, 16))
        gbs = nbytes / (ms * 1e-3) / 1e9
        out["row_mapped"][k] = {"lanes_per_row": G, "us": round(1e3 * ms, 1), "GB/s": round(gbs, 1),
                                "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 3)}
    del h, A
    src = torch.empty((1 << 30) // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * (1 << 30) / (e0.elapsed_time(e1) * 1e-3) / 1e9
    out["device_copy_1GiB_GB/s"] = round(copy_gbs, 1)

    for v in out["kernels"].values():
        v["frac_of_copy"] = round(v["GB/s"] / copy_gbs, 3)
    return out


def assemble_timing(sysm, dev, reps=5):
    """hsc_assemble_bcoo on the 78k-tet system re-stated as general triplets
    (both triangles: every off-diagonal block mirrored, transposed) -> the
    plugin's upper BCOO; median wall time of `reps` calls (each ends with a
    stream synchronisation), and the check that the result is the input."""
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin
    off = sysm.rows != sysm.cols
    vt = sysm.vals[off].reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)
    R = np.concatenate([sysm.rows, sysm.cols[off]])
    C = np.concatenate([sysm.cols, sysm.rows[off]])
    V = np.concatenate([sysm.vals, vt])
    perm = np.random.default_rng(0).permutation(len(R))
    r = torch.as_tensor(R[perm], dtype=torch.int32, device=dev)
    c = torch.as_tensor(C[perm], dtype=torch.int32, device=dev)
    v = torch.as_tensor(V[perm], dtype=torch.float64, device=dev).contiguous()
    out = plugin.assemble_bcoo(r, c, v, sysm.N // 3, fold_lower=False)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = plugin.assemble_bcoo(r, c, v, sysm.N // 3, fold_lower=False)
        ts.append(time.perf_counter() - t0)
    order = np.lexsort((sysm.cols, sysm.rows))
    same = bool(np.array_equal(out.rows.cpu().numpy(), sysm.rows[order]) and
                np.array_equal(out.values.cpu().numpy(), sysm.vals[order]))
    ms = 1e3 * float(np.median(ts))
    # the oracle (NumPy: stable argsort + in-order sums) on the same triplets
    from oracle import assembly
    tc = []
    for _ in range(3):
        t0 = time.perf_counter()
        ro, _, vo = assembly.assemble_bcoo(R[perm], C[perm], V[perm], fold_lower=False)
        tc.append(time.perf_counter() - t0)
    same_cpu = bool(np.array_equal(out.values.cpu().numpy(), vo))
    # algorithmic bytes: the triplets read once (4 + 4 + 72 B), the unique blocks written once
    nbytes = 80 * len(R) + 80 * int(out.rows.numel())
    return {"input_blocks": int(len(R)), "output_blocks": int(out.rows.numel()), "ms": round(ms, 3),
            "blocks_per_s": round(len(R) / (ms * 1e-3), 1), "output_equals_plugin_input": same,
            "algorithmic_GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "bound": "sort (hipcub SortPairs, 64-bit keys: block sort + ~20 merge launches); not HBM",
            "cpu_oracle_ms": round(1e3 * float(np.median(tc)), 2), "matches_oracle_bitwise": same_cpu,
            "cpu_sample": "the whole input through oracle.assembly (NumPy stable argsort + in-order sums), one thread",
            "what": "general triplets (both triangles, shuffled) -> upper-triangle BCOO (ge2sym + convert)"}


def fem_generator_timing(dev, reps=5, sample=64):
    """hsc_fem_hessian + hsc_assemble_bcoo on the 78,400 tets of the vehicle
    system's 7 tori (11,200 Tet4 each, E 1e10 on 4 and 1e8 on 3, nu 0.45),
    nodes displaced by N(0, 2 mm) so that the SPD projection acts: median wall
    time of `reps` calls each; a sample of tets checked against the oracle
    (1e-10 per block), whose own time on that sample is the CPU figure."""
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin, systems
    from oracle import fem
    rng = np.random.default_rng(0)
    nodes0, tets0, _ = systems.torus_mesh(0.45, 0.12, 0.06, 70, 16, 2)
    youngs = (1e10,) * 4 + (1e8,) * 3
    nodes = np.concatenate([nodes0 + np.array([1.2 * t, 0.0, 0.0]) for t in range(7)])
    tets = np.concatenate([tets0 + t * len(nodes0) for t in range(7)])
    E = np.repeat(np.array(youngs), len(tets0))
    nu = 0.45
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    X = nodes[tets]
    Dm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1)
    Dinv = np.linalg.inv(Dm)
    vol = np.abs(np.linalg.det(Dm)) / 6
    x = nodes + 0.002 * rng.normal(size=nodes.shape)
    t = lambda a, d: torch.as_tensor(np.ascontiguousarray(a), dtype=d, device=dev)  # noqa: E731
    args = (t(tets, torch.int32), t(x, torch.float64), t(Dinv, torch.float64), t(mu, torch.float64),
            t(lam, torch.float64), t(vol, torch.float64))
    rows, cols, vals = plugin.fem_hessian(*args, dt=0.01)
    A = plugin.assemble_bcoo(rows, cols, vals, len(nodes))
    tg, ta = [], []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rows, cols, vals = plugin.fem_hessian(*args, dt=0.01)
        t1 = time.perf_counter()
        A = plugin.assemble_bcoo(rows, cols, vals, len(nodes))
        t2 = time.perf_counter()
        tg.append(t1 - t0)
        ta.append(t2 - t1)
    vh = vals.cpu().numpy()
    ok = True
    t0 = time.perf_counter()
    for e in rng.choice(len(tets), sample, replace=False):
        for o, (a, b, B) in enumerate(fem.tet_blocks(x[tets[e]], Dinv[e], mu[e], lam[e], vol[e], dt=0.01)):
            tet = tets[e]
            ref = B if tet[a] <= tet[b] else B.T
            ok &= bool(np.abs(vh[10 * e + o].reshape(3, 3, order="F") - ref).max() <= 1e-10 * np.abs(ref).max())
    cpu_per_tet = (time.perf_counter() - t0) / sample
    g_ms, a_ms = 1e3 * float(np.median(tg)), 1e3 * float(np.median(ta))
    return {"tets": int(len(tets)), "triplets": int(10 * len(tets)), "blocks_assembled": int(A.rows.numel()),
            "generate_ms": round(g_ms, 3), "assemble_ms": round(a_ms, 3),
            "tets_per_s": round(len(tets) / (g_ms * 1e-3), 1), "matches_oracle_sample": ok,
            "cpu_oracle_tets_per_s": round(1.0 / cpu_per_tet, 1),
            "cpu_sample": f"{sample} tets through the oracle (pure Python Jacobi), one thread"}


def coupling_timing(dev, reps=7, sample=256):
    """hsc_abd_fem_coupling on random contact blocks between 12 bodies of
    2,000 ABD vertices and the vehicle system's 23,520 FEM nodes: 8,400
    blocks (its 2,100 coupled rim nodes x 4 ABD vertices each -> 33,600
    triplets) and 1M blocks (the kernel's bandwidth: 4 + 4 + 72 + 4 + 24 +
    4 + 4 bytes read and 4 x 80 written per contact, 436 B), median wall time
    of `reps` calls through the C ABI; a sample of contacts checked against
    the oracle (bitwise), whose own time on it is the CPU figure."""
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin
    from oracle import coupling
    rng = np.random.default_rng(0)
    nb, vpb, nf = 12, 2000, 7 * 3360
    v2b = np.repeat(np.arange(nb), vpb).astype(np.int32)
    xbar = rng.uniform(-0.3, 0.3, size=(nb * vpb, 3))
    body_fixed = np.zeros(nb, dtype=np.int32)
    vertex_fixed = np.zeros(nf, dtype=np.int32)
    t = lambda a, d: torch.as_tensor(np.ascontiguousarray(a), dtype=d, device=dev)  # noqa: E731
    common = (t(v2b, torch.int32), t(xbar, torch.float64), t(body_fixed, torch.int32), t(vertex_fixed, torch.int32),
              0, 4 * nb)
    out = {}
    for name, n in (("vehicle", 8400), ("1M", 1 << 20)):
        av = rng.integers(0, nb * vpb, n).astype(np.int32)
        fv = rng.integers(0, nf, n).astype(np.int32)
        H = rng.normal(size=(n, 3, 3))
        args = (t(av, torch.int32), t(fv, torch.int32), t(H, torch.float64)) + common
        plugin.abd_fem_coupling(*args)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rows, cols, vals = plugin.abd_fem_coupling(*args)
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(ts))
        pick = rng.choice(n, sample, replace=False)
        t0 = time.perf_counter()
        r, c, v = coupling.coupling_blocks(av[pick], fv[pick], H[pick], v2b, xbar, body_fixed, vertex_fixed, 0,
                                           4 * nb)
        cpu = (time.perf_counter() - t0) / sample
        idx = (4 * pick[:, None] + np.arange(4)[None]).ravel()
        ok = bool(np.array_equal(vals[torch.as_tensor(idx, device=dev)].cpu().numpy(), v)
                  and np.array_equal(rows[torch.as_tensor(idx, device=dev)].cpu().numpy(), r))
        out[name] = {"contacts": n, "triplets": 4 * n, "ms": round(ms, 4),
                     "contacts_per_s": round(n / (ms * 1e-3), 1), "GB_per_s_wall": round(436 * n / (ms * 1e-3) / 1e9, 1),
                     "matches_oracle_sample": ok, "cpu_oracle_contacts_per_s": round(1.0 / cpu, 1)}
    out["cpu_sample"] = f"{sample} contacts per case through the oracle (NumPy per contact), one thread"
    return out


def newton_replay(sysm, dev):
    """The Newton step's use of the direction (hsc_reimpl_amd.replay) on the
    78k-tet system: plain FP64 and the new solver at rtol 1e-8 and at
    libuipc's default tol_rate 1e-3 -- descent, full residual, the model line
    search's alpha and the model decrease as a fraction of the exact Newton
    step's.  Exact step: a plain FP64 solve at rtol 1e-12 (SciPy's direct
    solve if that does not converge)."""
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin, replay
    A = plugin.BCOO3.from_system(sysm, dev)
    b = torch.as_tensor(sysm.b, device=dev)
    h = plugin.Handle(plugin.make_config({"linear_system/tol_rate": 1e-12, "linear_system/hsc/max_iter": 20000},
                                         abd_inverse="exact"))
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    x = torch.zeros_like(b)
    st = h.solve(b, x, check=False)
    exact = {"how": "plain FP64 solve, rtol 1e-12", "status": st["status_name"],
             "full_relative_residual": st["full_relative_residual"]}
    x_ref = x.cpu().numpy()
    if st["status"] != 0:
        import scipy.sparse.linalg as spla
        x_ref = spla.spsolve(sysm.sparse().tocsc(), np.asarray(sysm.b))
        exact = {"how": "SciPy spsolve (the rtol 1e-12 solve ended with " + st["status_name"] + ")"}
    keep = ("status", "pcg_iterations", "descent_ok", "residual_ok", "full_relative_residual", "alpha", "trials",
            "accepted", "decrease_vs_exact")
    rows = {}
    for tol in (1e-8, 1e-3):
        for r in replay.replay(sysm, ["plain_fp64", "new"], {"linear_system/tol_rate": tol}, dev, x_exact=x_ref):
            rows[f"{r['variant']} @ tol {tol:g}"] = {k: r[k] for k in keep}
    return {"exact_step": exact, "directions": rows,
            "not_replayed": "feasible-step / CCD / CFL caps and the true energy (scene geometry)"}


def _sparse_solves(sysm, order, omega, rtol, budget_s, max_solves):
    from oracle import sparse as osp
    import numpy as np
    n = sysm.fn
    setup, solve, iters = [], [], 0
    while len(setup) < max_solves and sum(setup) + sum(solve) < budget_s:
        t0 = time.perf_counter()
        ref = osp.SparseHSC.from_system(sysm, order=order, omega=omega)
        t1 = time.perf_counter()
        _, _, st = ref.solve(sysm.b[:n], sysm.b[n:], rtol=rtol, maxiter=20000)
        t2 = time.perf_counter()
        setup.append(t1 - t0)
        solve.append(t2 - t1)
        iters = st["iterations"]
    tot = [a + b for a, b in zip(setup, solve)]
    return float(np.median(tot)), float(np.median(setup)), float(np.median(solve)), len(tot), sum(tot), iters


def _dense_c1(reps=5):
    """The dense restatement (core.py's HSC: dense D, the Cholesky check of
    the Schur complement, exact A^-1 as the plugin path) on the 2,178-DOF
    system: ACA tol 1e-6, Neumann order 2, omega 0.6 (0.7 makes the order-2
    series indefinite on it), rtol 1e-8; construction
    and solve timed separately, median of `reps` after one warm-up."""
    from hsc_reimpl_amd import systems
    from oracle import hsc_core
    import numpy as np
    s = systems.cube_system(seed=0)
    H = s.dense()
    n = s.fn
    D, A, E = H[:n, :n], H[n:, n:], H[n:, :n]
    setup, solve, iters = [], [], 0
    for r in range(reps + 1):
        t0 = time.perf_counter()
        hsc = hsc_core.HSC(D, A, E, order=2, omega=0.6, aca_rtol=1e-6, exact_affine=True)
        t1 = time.perf_counter()
        _, st = hsc.solve(s.b, rtol=1e-8, maxiter=5000)
        t2 = time.perf_counter()
        if r:
            setup.append(t1 - t0)
            solve.append(t2 - t1)
        iters = st["iterations"]
    return {"threads": int(os.environ.get("OPENBLAS_NUM_THREADS", "0")),
            "construct_s": round(float(np.median(setup)), 4),
            "solve_s": round(float(np.median(solve)), 4), "pcg_iterations": iters,
            "solves_per_s": round(1.0 / (float(np.median(setup)) + float(np.median(solve))), 3)}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), else the machine type."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_all_cores(order, omega, rtol, budget_s=5.0, max_solves=3):
    """The CPU measurements with the BLAS / OpenMP pools at the process's
    thread setting (a child process started with OMP_NUM_THREADS,
    OPENBLAS_NUM_THREADS and MKL_NUM_THREADS set before NumPy loads): the
    sparse restatement on the 78k-tet system and the dense one on the
    2,178-DOF system.  CPU only; nothing here touches the GPU."""
    from hsc_reimpl_amd import systems
    sysm = systems.torus_system(seed=1)
    med, m_setup, m_solve, k, _, _ = _sparse_solves(sysm, order, omega, rtol, budget_s, max_solves)
    return {"threads": int(os.environ.get("OPENBLAS_NUM_THREADS", "0")), "solves_per_s": round(1.0 / med, 4),
            "setup_s": round(m_setup, 3), "solve_s": round(m_solve, 3), "solves": k,
            "note": "SciPy's sparse products run on one thread whatever the BLAS pool",
            "dense_2178_dof": _dense_c1()}


def cpu_baseline(sysm, order, omega, rtol, budget_s=10.0, max_solves=8):
    """Oracle (sparse NumPy/SciPy restatement of the reference algorithm):
    full solves (setup + PCG + back-substitution) of the same system, single
    thread, repeated until about `budget_s` of CPU work; median rate (the
    reported value).  Beside it: the dense core.py-form restatement on the
    2,178-DOF system at one thread, and both again with the thread pools at
    the box's CPU share, in a child process whose pools are set before NumPy
    loads (this process keeps its single-thread pools)."""
    import subprocess
    import numpy as np
    med, m_setup, m_solve, k, tot, iters = _sparse_solves(sysm, order, omega, rtol, budget_s, max_solves)
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": 1, "kind": "port",
           "sample": f"{k} full solves (setup + {iters} PCG iterations each) of the same 78k-tet system, "
                     f"median {med:.2f} s, {tot:.1f} s total; SciPy sparse matvecs, BLAS pinned to 1 thread",
           "iterations": iters, "cpu": cpu_model(), "nproc": os.cpu_count(),
           "setup_s": round(m_setup, 3), "solve_s": round(m_solve, 3),
           "dense_2178_dof_one_thread": dict(_dense_c1(), what="core.py-form dense restatement "
                                                               "(oracle.hsc_core.HSC) on the 2,178-DOF system")}
    allc = max(1, min(os.cpu_count() or 1, 16))
    env = dict(os.environ, OMP_NUM_THREADS=str(allc), OPENBLAS_NUM_THREADS=str(allc), MKL_NUM_THREADS=str(allc))
    try:
        res = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-only"], env=env,
                             capture_output=True, text=True, timeout=240)
        line = [x for x in res.stdout.splitlines() if x.startswith("{")]
        out["all_cores"] = json.loads(line[-1]) if res.returncode == 0 and line else {
            "threads": allc, "error": f"exit {res.returncode}: {res.stderr[-300:]}"}
    except subprocess.TimeoutExpired:
        out["all_cores"] = {"threads": allc, "error": "timed out after 240 s"}
    return out


def run_batch64(args, dev, order, omega, rtol):
    """64 uneven systems (seed 2: 1-16 ABD bodies, 1-8 tori of 2k-15k tets,
    ranks 4-16), the new solver (FP32 bulk / FP64 reductions, 16 Chebyshev-
    weighted passes, ratio 1000; one interval for the batch), rtol 1e-8 each: one captured
    hipGraph replay (right-hand sides, one persistent PCG launch over all 64,
    back-substitution and acceptance) vs the same systems solved one at a
    time.  A step = one solve of all 64 systems (setup excluded: the systems
    are set up once, as a batch of Newton systems would be between solves)."""
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin, systems

    t0 = time.perf_counter()
    syss = systems.batch_systems(seed=2, workers=min(16, os.cpu_count() or 1))
    print(f"[batch64] generated 64 systems ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    scene = {"linear_system/tol_rate": rtol, "linear_system/hsc/neumann_order": order,
             "linear_system/hsc/omega": omega, "linear_system/hsc/max_iter": 20000}
    cfg = dict(precision="mixed", residual_replace_every=25, preconditioner="chebyshev",
               poly_interval_ratio=NEW_RATIO, abd_inverse="exact")
    hs, As, bs = [], [], []
    t0 = time.perf_counter()
    for s in syss:
        h = plugin.Handle(plugin.make_config(scene, **cfg))
        A = plugin.BCOO3.from_system(s, dev)
        h.setup(A, s.qo, s.qm, s.fo, s.fn)
        hs.append(h)
        As.append(A)
        bs.append(torch.as_tensor(s.b, device=dev))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"[batch64] set up 64 systems in {setup_s:.1f} s", file=sys.stderr, flush=True)
    # one at a time
    xs = [torch.zeros_like(b) for b in bs]
    for h, b, x in zip(hs, bs, xs):
        h.solve(b, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ind = [h.solve(b, x) for h, b, x in zip(hs, bs, xs)]
    torch.cuda.synchronize()
    t_ind = (time.perf_counter() - t0) / args.steps
    print(f"[batch64] one at a time: {1e3 * t_ind:.1f} ms per pass", file=sys.stderr, flush=True)
    # captured batch
    xg = [torch.zeros_like(b) for b in bs]
    batch = plugin.Batch(hs, bs, xg)
    for _ in range(args.warmup):
        batch.launch()
        batch.stats()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    replay_ms = []
    for _ in range(args.steps):
        e0.record()
        batch.launch()
        e1.record()
        st = batch.stats()
        replay_ms.append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / args.steps
    ok = all(d["status"] == 0 for d in st) and all(d["status"] == 0 for d in ind)
    diff = max(float((torch.linalg.vector_norm(a - c) / torch.linalg.vector_norm(c)).item()) for a, c in zip(xg, xs))
    byts = 0
    for h, d in zip(hs, st):
        i = h.setup_info()
        byts += pcg_bytes_per_iteration(i["fem_nodes"], i["nnzb"], i["active_fem_dofs"], i["rank"], order,
                                        mixed=True) * d["pcg_iterations"]
    iters = [d["pcg_iterations"] for d in st]
    med_ms = float(np.median(replay_ms))
    achieved = byts / (med_ms * 1e-3) / 1e9
    line = {
        "metric": "batched HSC: systems/s @ rtol 1e-8, 64 uneven systems (seed 2), captured batch vs one at a time",
        "value": round(64 / t_b, 2), "unit": "systems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_b, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic (seeded generator; no external data)",
        "config": {"workload": "64 uneven HSC systems, one captured persistent launch",
                   "tets": int(sum(s.info["tets"] for s in syss)), "dofs": int(sum(s.N for s in syss)),
                   "nnzb": int(sum(h.setup_info()["nnzb"] for h in hs)), "neumann_order": order, "omega": omega,
                   "rtol": rtol, **batch.info()},
        "one_at_a_time": {"systems_per_s": round(64 / t_ind, 2), "ms_per_pass": round(1e3 * t_ind, 3)},
        "speedup_vs_one_at_a_time": round(t_ind / t_b, 3),
        "all_converged": ok, "max_rel_diff_vs_one_at_a_time": diff,
        "pcg_iterations": {"min": int(min(iters)), "median": int(np.median(iters)), "max": int(max(iters))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "captured batch replay (rhs + pcg_persistent<G, float> over 64 systems + finish)",
                     "bytes_per_replay": int(byts), "replay_ms": round(med_ms, 3)},
        "setup_s_all_64": round(setup_s, 2),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one mixed solve with the launch shape pinned (no tuner launches), then a 256 MiB "
                         "device copy: the program profiled by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--bandwidth-probe", action="store_true",
                    help="one launch of each streaming operator kernel on 32 copies of the 78k-tet system's "
                         "FEM part, then a 256 MiB device copy: the program profiled by rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE for the kernels' traffic")
    ap.add_argument("--workload", choices=["78k-tet", "batch64"], default="78k-tet",
                    help="78k-tet: the headline single-system solve; batch64: 64 uneven systems (seed 2) "
                         "in one captured persistent launch vs one at a time")
    ap.add_argument("--lanes-per-row", type=int, default=4)
    ap.add_argument("--wg-per-cu", type=int, default=1)
    ap.add_argument("--schedule", type=int, default=3)
    ap.add_argument("--preconditioner", default="learned", choices=["neumann", "chebyshev", "learned"],
                    help="preconditioner of the --traffic-probe solve")
    ap.add_argument("--cpu-only", action="store_true",
                    help="print the CPU baseline's measurements at the inherited thread setting and exit "
                         "(the default run starts this as a child process; no GPU use)")
    args = ap.parse_args()
    if args.cpu_only:
        print(json.dumps(cpu_all_cores(5, 0.7, 1e-8)), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: the HSC solve of one system runs on one MI355X and no "
              "multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline is a single-thread measurement: pin BLAS before NumPy loads
    for var in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[var] = "1"
    import numpy as np
    import torch
    from hsc_reimpl_amd import plugin, systems

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    order, omega, rtol = 5, 0.7, 1e-8
    if args.workload == "batch64":
        return run_batch64(args, dev, NEW_PASSES, omega, rtol)
    sysm = systems.torus_system(seed=1)
    A = plugin.BCOO3.from_system(sysm, dev)
    b = torch.as_tensor(sysm.b, device=dev)
    scene = {"linear_system/tol_rate": rtol, "linear_system/hsc/neumann_order": order,
             "linear_system/hsc/omega": omega, "linear_system/hsc/max_iter": 20000}

    new_scene = dict(scene, **{"linear_system/hsc/neumann_order": NEW_PASSES})
    if args.bandwidth_probe:
        big = systems.replicate_fem(sysm, 32)
        Ab = plugin.BCOO3.from_system(big, dev)
        h = plugin.Handle(plugin.make_config({"linear_system/tol_rate": rtol}, precision="mixed"))
        h.setup(Ab, big.qo, big.qm, big.fo, big.fn)
        i = h.setup_info()
        nb, nnzb = i["fem_nodes"], i["nnzb"]
        ms = {k: h.time_kernel(k, 0, repeats=1) for k in ("spmv64", "pass64", "pass32")}
        nbytes = 256 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        n = 3 * nb
        alg = {"spmv64": 76 * nnzb + 4 * (nb + 1) + 16 * n, "pass64": 76 * nnzb + 4 * (nb + 1) + 24 * n + 72 * nb,
               "pass32": 40 * nnzb + 4 * (nb + 1) + 16 * n + 36 * nb}
        print(json.dumps({"probe": "bandwidth", "fem_nodes": nb, "nnzb": nnzb, "algorithmic_bytes": alg,
                          "ms_single_launch": ms, "copy_bytes": nbytes}), flush=True)
        return
    if args.traffic_probe:
        poly = args.preconditioner != "neumann"
        ext = {"poly_interval_ratio": NEW_RATIO} if args.preconditioner == "chebyshev" else {}
        h = plugin.Handle(plugin.make_config(new_scene if poly else scene, precision="mixed",
                                             residual_replace_every=25, lanes_per_row=args.lanes_per_row,
                                             wg_per_cu=args.wg_per_cu, schedule=args.schedule,
                                             preconditioner=args.preconditioner,
                                             abd_inverse="exact" if poly else "neumann", **ext))
        if args.preconditioner == "learned":
            from hsc_reimpl_amd import learned
            w = torch.as_tensor(learned.load_weights(), device=dev)
            h._check(h.lib.hsc_set_learned_weights(h.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
        h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
        x = torch.zeros_like(b)
        st = h.solve(b, x)
        torch.cuda.synchronize()
        nbytes = 256 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        info = h.setup_info()
        print(json.dumps({"probe": "traffic", "pcg_iterations": st["pcg_iterations"], "pcg_ms": st["pcg_ms"],
                          "bytes_per_iteration": pcg_bytes_per_iteration(info["fem_nodes"], info["nnzb"],
                                                                         info["active_fem_dofs"], info["rank"],
                                                                         NEW_PASSES if poly else order,
                                                                         mixed=True),
                          "copy_bytes": nbytes}), flush=True)
        return

    from hsc_reimpl_amd import learned
    theta = learned.load_weights()

    def run(precision, preconditioner="neumann"):
        ext = {"precision": "mixed", "residual_replace_every": 25} if precision == "mixed" else {}
        ext["preconditioner"] = preconditioner
        if preconditioner != "neumann":
            # the polynomial variants run with the exact ABD inverse; the Neumann
            # variants keep the plugin's default (block Neumann + doubling)
            ext["abd_inverse"] = "exact"
        if preconditioner == "chebyshev":
            ext["poly_interval_ratio"] = NEW_RATIO
        h = plugin.Handle(plugin.make_config(scene if preconditioner == "neumann" else new_scene, **ext))
        if preconditioner == "learned":
            w = torch.as_tensor(theta, device=dev)
            h._check(h.lib.hsc_set_learned_weights(h.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
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

    plain = run("fp64")                      # the reference algorithm: Neumann series, FP64
    mixed = run("mixed")                     # + FP32 bulk / FP64 reductions
    new = run("mixed", "learned")            # + 16 passes on the network's interval
    cheb = run("mixed", "chebyshev")         # the same passes on the fixed interval ratio 1000
    cheb64 = run("fp64", "chebyshev")
    info = plain["info"]
    info_new = new["info"]
    nb, nnzb, a, k = info["fem_nodes"], info["nnzb"], info["active_fem_dofs"], info["rank"]
    err_vs_fp64 = float(np.linalg.norm(new["x"] - plain["x"]) / np.linalg.norm(plain["x"]))
    # accuracy against a sparse direct solve of the full system (SciPy SuperLU):
    # the plugin's default ABD inverse is accurate to abd_inverse_tol (1e-5),
    # the new solver's exact one to the PCG tolerance
    import scipy.sparse.linalg as spla
    x_direct = spla.spsolve(sysm.sparse().tocsc(), np.asarray(sysm.b))
    err_direct = {name: float(np.linalg.norm(r["x"] - x_direct) / np.linalg.norm(x_direct))
                  for name, r in (("plain_fp64", plain), ("new", new), ("mixed_chebyshev_1000", cheb))}
    bpi = pcg_bytes_per_iteration(nb, nnzb, a, k, NEW_PASSES, mixed=True)
    achieved = bpi * new["pcg_iterations"] / (new["pcg_ms"] * 1e-3) / 1e9
    bpi64 = pcg_bytes_per_iteration(nb, nnzb, a, k, order, mixed=False)
    achieved64 = bpi64 * plain["pcg_iterations"] / (plain["pcg_ms"] * 1e-3) / 1e9
    traffic = None
    prof = ROOT / "profiles" / "pcg_persistent_traffic.json"
    if prof.exists():
        traffic = json.loads(prof.read_text()).get("traffic_bytes_per_launch_new")
    # mixed vs FP64 on the 2,178-DOF system across FEM stiffness (the mixed-
    # precision configuration's sweep); error of the mixed solution vs the FP64 one
    sweep = []
    c1_scene = {"linear_system/tol_rate": rtol, "linear_system/hsc/neumann_order": 2,
                "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 20000}
    for E in (1e6, 1e8, 1e10):
        sc1 = systems.cube_system(seed=0, E=E)
        A1 = plugin.BCOO3.from_system(sc1, dev)
        b1 = torch.as_tensor(sc1.b, device=dev)
        res = {}
        for prec in ("fp64", "mixed"):
            ext = {"aca_rtol": 1e-6}
            if prec == "mixed":
                ext.update(precision="mixed", residual_replace_every=25)
            h1 = plugin.Handle(plugin.make_config(c1_scene, **ext))
            h1.setup(A1, sc1.qo, sc1.qm, sc1.fo, sc1.fn)
            x1 = torch.zeros_like(b1)
            st1 = h1.solve(b1, x1)
            res[prec] = (st1, x1)
        diff = float((torch.linalg.vector_norm(res["mixed"][1] - res["fp64"][1]) /
                      torch.linalg.vector_norm(res["fp64"][1])).item())
        sweep.append({"E": E, "iterations": {"fp64": res["fp64"][0]["pcg_iterations"],
                                             "mixed": res["mixed"][0]["pcg_iterations"]},
                      "mixed_rel_diff_vs_fp64": diff,
                      "schur_relative_residual_mixed": res["mixed"][0]["schur_relative_residual"]})
    # what the network gains on the 2,178-DOF system and a held-out one of its
    # family: iterations and device ms beside the plain Neumann series and the
    # fixed intervals
    rows = learned.evaluate(theta, {"c1": systems.cube_system(seed=0), "cube200": learned.family_system(200)},
                            device=dev, reps=3)
    learned_c1 = {"systems": "2,178-DOF system and seed 200 of its family (Neumann order 2, omega 0.6; "
                             "polynomials at 16 passes; mixed precision, rtol 1e-8)",
                  "rows": [{"system": r["system"],
                            "iterations": {k: r[k]["iterations"] for k in
                                           ("neumann", "chebyshev_100", "chebyshev_1000", "learned")},
                            "pcg_ms": {k: round(r[k]["pcg_ms"], 3) for k in
                                       ("neumann", "chebyshev_100", "chebyshev_1000", "learned")},
                            "learned_ratio": round(r["learned"]["ratio"], 1)} for r in rows]}
    bandwidth = beyond_mall(sysm, dev)
    assembly = assemble_timing(sysm, dev)
    generator = fem_generator_timing(dev)
    coupling_t = coupling_timing(dev)
    replay_r = newton_replay(sysm, dev)
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
                   "new_passes": NEW_PASSES, "new_interval_ratio": round(info_new["poly_ratio"], 1),
                   "new_interval_from": "learned graph MLP", "pass_sync": info_new["pass_sync"],
                   "pass_pairs": info_new["pass_pairs"],
                   "lanes_per_row": info_new["lanes_per_row"], "wg_per_cu": info_new["workgroups"],
                   "schedule": info_new["schedule"]},
        "new": "learned + mixed precision: FP32 bulk, FP64 reductions and true residual every 25; 16 block-"
               "Jacobi passes weighted by the Chebyshev roots of [b/ratio, b] in Leja order, the ratio picked at "
               "setup by the graph MLP from D's node features, b = 1.2 x a power-iteration bound; passes in pairs: "
               "each row tile gathers its node set once, computes its own and halo rows' pass in LDS, then the next "
               "pass; between pairs the tiles exchange tagged pass values (value and round in one atomic word) "
               "instead of waiting on flags",
        "pcg_iterations": {"new": new["pcg_iterations"], "plain_fp64": plain["pcg_iterations"],
                           "mixed_chebyshev_1000": cheb["pcg_iterations"],
                           "mixed_neumann": mixed["pcg_iterations"], "fp64_chebyshev": cheb64["pcg_iterations"]},
        "plain_fp64": {"solves_per_s": round(plain["solves_per_s"], 3), "ms_per_solve": round(plain["ms_per_solve"], 3),
                       "pcg_ms": round(plain["pcg_ms"], 3), "pcg_gbs": round(achieved64, 1)},
        "mixed_chebyshev_1000": {"solves_per_s": round(cheb["solves_per_s"], 3),
                                 "ms_per_solve": round(cheb["ms_per_solve"], 3), "pcg_ms": round(cheb["pcg_ms"], 3)},
        "mixed_neumann": {"solves_per_s": round(mixed["solves_per_s"], 3),
                          "ms_per_solve": round(mixed["ms_per_solve"], 3), "pcg_ms": round(mixed["pcg_ms"], 3)},
        "fp64_chebyshev": {"solves_per_s": round(cheb64["solves_per_s"], 3),
                           "ms_per_solve": round(cheb64["ms_per_solve"], 3), "pcg_ms": round(cheb64["pcg_ms"], 3)},
        "speedup_vs_plain_fp64": round(new["solves_per_s"] / plain["solves_per_s"], 3),
        "fp64_solution_rel_diff": err_vs_fp64,
        "solution_rel_error_vs_direct": err_direct,
        "schur_relative_residual": new["schur_relative_residual"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "pcg_persistent<G, float, true> (whole Schur PCG of the new solver, one cooperative launch)",
                     "kernel_name_in_trace": "void hsc::pcg_persistent<%d, float, true>(hsc::PcgParams)"
                                             % info_new["lanes_per_row"],
                     "bytes_per_iteration": bpi, "iterations": new["pcg_iterations"],
                     "kernel_ms": round(new["pcg_ms"], 3)},
        "cpu_baseline": cpu,
        "learned_preconditioner_2178_dof": learned_c1,
        "mixed_stiffness_sweep_2178_dof": sweep,
        "bandwidth_beyond_mall": bandwidth,
        "assembly_before_path": assembly,
        "fem_hessian_generator": generator,
        "abd_fem_coupling": coupling_t,
        "newton_replay": replay_r,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
