"""Replay of the Newton step's use of the solver's direction on saved systems
(the caller side after the path: advance_ipc.cu:376-440 of the reference,
with the plugin's acceptance checks of linear_hsc.cu:400-435).

A saved Newton system is H dx = b, where b = -grad E at the start point x0.
Each solver variant's direction dx is checked the way the step uses it:

- descent: b^T dx > 0 (linear_hsc.cu:400-402);
- full residual: |b - H dx| / |b| <= full_residual_limit
  (linear_hsc.cu:433-435);
- line search: alpha = 1, halved until the energy does not increase
  (advance_ipc.cu:410-440). Here the energy is the quadratic model at x0,
  E(alpha) - E0 = -alpha b^T dx + alpha^2 / 2 dx^T H dx.

The feasible-step, CCD and CFL caps (advance_ipc.cu:392-406) and the true
energy need the scene's geometry, which lives in libuipc and not in a saved
system. They are not replayed, so alpha starts at 1. The model decrease at
the accepted alpha is reported as a fraction of the exact Newton step's
decrease, -b^T dx* / 2 (dx* = H^-1 b by a sparse direct solve).

The solves run on the GPU through the C ABI. The model evaluation (one
sparse product H dx per variant) runs on the host.

    python -m hsc_reimpl_amd.replay A.mtx b.mtx --qo 0 --qm 144 --fo 144 --fn 70560
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

VARIANTS = {
    # the plugin's algorithm with its defaults (block Neumann ABD inverse, doubled)
    "plain_fp64": {},
    "mixed_neumann": {"precision": "mixed"},
    # the new solver: mixed precision, 16 passes on the interval the committed
    # network picks, exact ABD inverse
    "new": {"precision": "mixed", "preconditioner": "learned", "neumann_order": 16, "abd_inverse": "exact"},
    "chebyshev_1000": {"precision": "mixed", "preconditioner": "chebyshev", "neumann_order": 16,
                       "poly_interval_ratio": 1000.0, "abd_inverse": "exact"},
}


def make_handle(scene: dict, extra: dict, device="cuda:0", **more):
    """A handle of one variant (make_config keys + extras); the learned
    preconditioner gets the committed network's weights."""
    from . import plugin
    h = plugin.Handle(plugin.make_config(scene, **extra, **more))
    if extra.get("preconditioner") == "learned":
        import ctypes as C

        import torch

        from . import learned
        w = torch.as_tensor(learned.load_weights(), device=device)
        h._check(h.lib.hsc_set_learned_weights(h.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
    return h


def model_line_search(bdx: float, dHd: float, max_iter: int = 8) -> dict:
    """The line search of advance_ipc.cu:410-440 on the quadratic model
    E(alpha) - E0 = -alpha bdx + alpha^2 dHd / 2: alpha = 1, halved after each
    trial whose energy exceeds E0, at most max_iter trials.  Returns the last
    alpha tried, the trial count, whether it was accepted and its model
    energy change."""
    alpha = 1.0
    de = 0.0
    for trial in range(1, max_iter + 1):
        de = -alpha * bdx + 0.5 * alpha * alpha * dHd
        if de <= 0.0:
            return {"alpha": alpha, "trials": trial, "accepted": True, "energy_change": de}
        if trial < max_iter:
            alpha /= 2.0
    return {"alpha": alpha, "trials": max_iter, "accepted": False, "energy_change": de}


def evaluate_direction(H, b: np.ndarray, dx: np.ndarray, full_residual_limit: float,
                       exact_decrease: float | None = None, max_iter: int = 8) -> dict:
    """Acceptance and model line search of one direction (host arithmetic)."""
    Hd = H @ dx
    bdx = float(b @ dx)
    dHd = float(dx @ Hd)
    rel = float(np.linalg.norm(b - Hd) / max(np.linalg.norm(b), 1e-300))
    ls = model_line_search(bdx, dHd, max_iter)
    out = {"descent": bdx, "descent_ok": bdx > 0.0, "full_relative_residual": rel,
           "residual_ok": rel <= full_residual_limit, **ls}
    if exact_decrease is not None and exact_decrease < 0.0:
        out["decrease_vs_exact"] = ls["energy_change"] / exact_decrease
    return out


def replay(sysm, variants=None, scene=None, device="cuda:0", x_exact=None, max_iter: int = 8) -> list:
    """Solve a saved system (systems.HSCSystem) with each variant through the
    C ABI and check the directions.  `variants`: names of VARIANTS or a dict
    name -> make_config extras; `scene`: libuipc-style keys (tol_rate, ...)."""
    import scipy.sparse.linalg as spla
    import torch

    from . import plugin
    if variants is None:
        variants = {k: VARIANTS[k] for k in ("plain_fp64", "new")}
    elif not isinstance(variants, dict):
        variants = {k: VARIANTS[k] for k in variants}
    scene = dict(scene or {})
    H = sysm.sparse()
    b = np.asarray(sysm.b, dtype=np.float64)
    if x_exact is None:
        x_exact = spla.spsolve(H.tocsc(), b)
    exact_decrease = -0.5 * float(b @ x_exact)
    A = plugin.BCOO3.from_system(sysm, device)
    bd = torch.as_tensor(b, device=device)
    rows = []
    for name, extra in variants.items():
        h = make_handle(scene, extra, device)
        cfg = h.config
        h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
        x = torch.zeros_like(bd)
        st = h.solve(bd, x, check=False)
        dx = x.cpu().numpy()
        row = {"variant": name, "status": st["status_name"], "pcg_iterations": st["pcg_iterations"],
               "pcg_ms": round(st["pcg_ms"], 3), "solver_accepted": bool(st["accepted"]),
               "solution_rel_diff": float(np.linalg.norm(dx - x_exact) / max(np.linalg.norm(x_exact), 1e-300))}
        row.update(evaluate_direction(H, b, dx, cfg.full_residual_limit, exact_decrease, max_iter))
        rows.append(row)
    return rows


def main(argv=None) -> int:
    from . import mtx, systems
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("A")
    ap.add_argument("b")
    for k in ("qo", "qm", "fo", "fn"):
        ap.add_argument(f"--{k}", type=int, required=True)
    ap.add_argument("--tol-rate", type=float, default=None)
    ap.add_argument("--variants", nargs="*", default=["plain_fp64", "new"], choices=sorted(VARIANTS))
    ap.add_argument("--line-search-max-iter", type=int, default=8)
    ap.add_argument("--device", default="cuda:0")
    args = ap.parse_args(argv)
    r, c, v, nbr = mtx.read_bcoo(args.A)
    b = mtx.read_vector(args.b)
    if 3 * nbr != len(b) or args.qm + args.fn != len(b):
        print(f"size mismatch: A has {3 * nbr} rows, b {len(b)}, qm + fn = {args.qm + args.fn}", file=sys.stderr)
        return 2
    sysm = systems.HSCSystem(r, c, v, b, fo=args.fo, fn=args.fn, qo=args.qo, qm=args.qm)
    scene = {} if args.tol_rate is None else {"linear_system/tol_rate": args.tol_rate}
    for row in replay(sysm, args.variants, scene, args.device, max_iter=args.line_search_max_iter):
        print(json.dumps(row), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
