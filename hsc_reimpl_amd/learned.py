"""Learned polynomial preconditioner (no reference counterpart).

The preconditioner keeps the reference's block-Jacobi passes over D with
per-pass weights from the Chebyshev roots of an interval [b/ratio, b]
(b = 1.2 x the setup's power-iteration estimate of lambda_max(Dinv D)): a
fixed SPD polynomial of Dinv D for the whole solve, so plain PCG runs in
the persistent kernel, in FP64 or mixed precision.  What is learned is the
interval: a per-node graph MLP over D's node graph (two message-passing
layers, hidden width 64, BF16 weights and activations, FP32 accumulation)
reads each node's features of D at setup, and the node mean of its three
outputs predicts log10(ratio) - 2.5 for degrees 8, 12 and 16.  The best
ratio depends on the system (about 300 for the 2,178-DOF family at degree
16, about 1000 for the 78k-tet tori); no single value serves both.

Everything numeric runs in the HIP library through the C ABI:
  * targets: each training system is solved (mixed precision, rtol 1e-8) at
    every ratio of RATIOS for each degree, and the ratio with the fewest PCG
    iterations (refined by a parabola through its neighbours in log ratio)
    is the target of every node of that system;
  * features: hsc_get_learned_features (the same kernel the setup runs);
  * training: hsc_mlp_train_step (hand-written BF16 forward/backward + Adam)
    on the disjoint union of the training systems' node graphs;
  * inference: hsc_set_learned_weights + preconditioner = "learned".
torch tensors are only the device buffers.

Train the committed weights with ``python -m hsc_reimpl_amd.learned``.
"""
from __future__ import annotations

import ctypes as C
import json
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _abi, plugin, systems

WEIGHTS = Path(__file__).resolve().parent / "data" / "learned_graph_mlp.npz"
# the 2,178-DOF family's reference settings (Neumann order 2, omega 0.6, rtol 1e-8)
C1_SCENE = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
            "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
# the 2,178-DOF family's extension settings: ACA stop 1e-6 (core.py's tol) and the
# exact ABD inverse, as in every solve the network was trained and evaluated on
C1_EXT = {"aca_rtol": 1e-6, "abd_inverse": "exact"}
DEGREES = (8, 12, 16)
RATIOS = (10.0, 20.0, 50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0, 5000.0, 10000.0, 20000.0)
CENTER = 2.5  # channel value = log10(ratio) - CENTER


def param_count() -> int:
    return int(_abi.load().hsc_mlp_param_count())


def init_params(seed=0) -> np.ndarray:
    """Flat float32 parameters in the library's layout (He init over the 14
    used inputs; output layer scaled by 0.1; padded rows/columns zero)."""
    rng = np.random.default_rng(seed)
    HID, IN, OUTP = 64, 16, 32
    parts = []
    W0 = rng.normal(size=(IN, HID)) * np.sqrt(2.0 / 14)
    W0[14:] = 0.0
    parts += [W0, np.zeros(HID)]
    for _ in range(2):
        parts += [rng.normal(size=(HID, HID)) * np.sqrt(2.0 / HID), rng.normal(size=(HID, HID)) * np.sqrt(2.0 / HID),
                  np.zeros(HID)]
    Wo = np.zeros((HID, OUTP))
    Wo[:, :3] = rng.normal(size=(HID, 3)) * np.sqrt(2.0 / HID) * 0.1
    parts += [Wo, np.zeros(OUTP)]
    theta = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
    assert theta.size == param_count()
    return theta


# flat parameter layout of the library (row-major matrices, Wo/bo padded to 32 outputs)
MLP_SHAPES = [("W0", (16, 64)), ("b0", (64,)), ("Ws1", (64, 64)), ("Wn1", (64, 64)), ("b1", (64,)),
              ("Ws2", (64, 64)), ("Wn2", (64, 64)), ("b2", (64,)), ("Wo", (64, 32)), ("bo", (32,))]
MLP_OFFSETS = {}
_o = 0
for _name, _shp in MLP_SHAPES:
    MLP_OFFSETS[_name] = _o
    _o += int(np.prod(_shp))
MLP_COUNT = _o


def to_oracle_params(theta: np.ndarray) -> dict:
    """Unpack the flat layout into the oracle's named parameters (Wo/bo
    trimmed to 3 outputs)."""
    out = {}
    for name, shp in MLP_SHAPES:
        o = MLP_OFFSETS[name]
        out[name] = np.asarray(theta[o:o + int(np.prod(shp))], dtype=np.float64).reshape(shp)
    out["Wo"] = out["Wo"][:, :3]
    out["bo"] = out["bo"][:3]
    return out


class GraphMLP:
    """The network object of the library (hsc_mlp_*) for one node graph."""

    def __init__(self, nbr_ptr: torch.Tensor, nbr: torch.Tensor, max_batch: int, device=None):
        self.lib = _abi.load()
        dev = torch.cuda.current_device() if device is None else device
        self.nb = int(nbr_ptr.numel()) - 1
        self.max_batch = max_batch
        h = C.c_void_p()
        code = self.lib.hsc_mlp_create(dev, self.nb, C.c_void_p(nbr_ptr.data_ptr()), C.c_void_p(nbr.data_ptr()),
                                       int(nbr.numel()), max_batch, C.byref(h))
        if code:
            raise plugin.HSCError(code, "hsc_mlp_create")
        self.h = h
        plugin.register_live("mlp", self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.hsc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta: np.ndarray):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        self._check(self.lib.hsc_mlp_set_params(self.h, t.ctypes.data_as(C.c_void_p), 0))

    def get_params(self) -> np.ndarray:
        t = np.zeros(param_count(), dtype=np.float32)
        self._check(self.lib.hsc_mlp_get_params(self.h, t.ctypes.data_as(C.c_void_p)))
        return t

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        B = X.shape[0]
        O = torch.zeros((B, self.nb, 3), dtype=torch.float32, device=X.device)
        self._check(self.lib.hsc_mlp_forward(self.h, C.c_void_p(X.data_ptr()), B, C.c_void_p(O.data_ptr()),
                                             plugin._stream(None)))
        return O

    def gradients(self, X, T, den):
        g = np.zeros(param_count(), dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.hsc_mlp_gradients(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(T.data_ptr()),
                                               C.c_void_p(den.data_ptr()), X.shape[0], g.ctypes.data_as(C.c_void_p),
                                               C.byref(loss), plugin._stream(None)))
        return float(loss.value), g

    def train_step(self, X, T, den, lr: float) -> float:
        loss = C.c_float()
        self._check(self.lib.hsc_mlp_train_step(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(T.data_ptr()),
                                                C.c_void_p(den.data_ptr()), X.shape[0], float(lr), C.byref(loss),
                                                plugin._stream(None)))
        return float(loss.value)

    def _check(self, code):
        if code:
            raise plugin.HSCError(code, "graph MLP")


def node_graph(handle: plugin.Handle, device="cuda"):
    ne = C.c_int64()
    lib = handle.lib
    handle._check(lib.hsc_get_node_graph(handle.h, None, None, C.byref(ne)))
    info = handle.setup_info()
    ptr = torch.zeros(info["fem_nodes"] + 1, dtype=torch.int32, device=device)
    nbr = torch.zeros(max(ne.value, 1), dtype=torch.int32, device=device)
    handle._check(lib.hsc_get_node_graph(handle.h, C.c_void_p(ptr.data_ptr()), C.c_void_p(nbr.data_ptr()),
                                         C.byref(ne)))
    return ptr, nbr[:ne.value]


def node_features(handle: plugin.Handle, device="cuda") -> torch.Tensor:
    """The network's input X [nb, 16] (bfloat16) of a set-up handle."""
    nb = handle.setup_info()["fem_nodes"]
    X = torch.zeros((nb, 16), dtype=torch.bfloat16, device=device)
    handle._check(handle.lib.hsc_get_learned_features(handle.h, C.c_void_p(X.data_ptr())))
    return X


def family_system(seed: int):
    """A system of the 2,178-DOF family with E log-uniform in [1e6, 1e10]."""
    E = float(np.exp(np.random.default_rng(10_000 + seed).uniform(np.log(1e6), np.log(1e10))))
    return systems.cube_system(seed=seed, E=E)


def torus_family_system(seed: int):
    """A small tire-like system: 1-4 ABD bodies, 1-2 tori of 2k-8k tets
    (log-uniform), E log-uniform in [1e6, 1e10], coupling rank 4-12."""
    rng = np.random.default_rng(30_000 + seed)
    ntor = int(rng.integers(1, 3))
    spec = dict(index=seed, bodies=int(rng.integers(1, 5)), tori=ntor, rank=int(rng.integers(4, 13)),
                tets=np.exp(rng.uniform(np.log(2000), np.log(8000), ntor)),
                youngs=np.exp(rng.uniform(np.log(1e6), np.log(1e10), ntor)), seed=int(rng.integers(1 << 31)))
    return systems.batch_system(spec)


def vehicle_family_system(seed: int):
    """A tire system at the 78k-tet system's resolution (16 cells around the
    tube, 2 through the wall): 1-3 tori of 50-90 cells around the ring, ABD
    hubs coupled at rank 8 over 200-300 rim nodes, E log-uniform in
    [1e8, 1e10]."""
    rng = np.random.default_rng(40_000 + seed)
    nt = int(rng.integers(1, 4))
    return systems.torus_system(seed=int(rng.integers(1 << 31)), n_tori=nt, n_bodies=nt,
                                youngs=tuple(np.exp(rng.uniform(np.log(1e8), np.log(1e10), nt))),
                                n_theta=2 * int(rng.integers(25, 46)), n_rim=int(rng.integers(200, 301)))


def training_system(key):
    """("cube" | "torus" | "vehicle", seed) -> the system (a picklable key,
    so systems are built in worker processes)."""
    kind, seed = key
    make = {"cube": family_system, "torus": torus_family_system, "vehicle": vehicle_family_system}[kind]
    return make(seed)


def poly_handle(sysm, degree, ratio=1000.0, precision="mixed", device="cuda", preconditioner="chebyshev",
                weights=None):
    """A set-up handle of the polynomial preconditioner (Chebyshev with the
    given ratio, or learned with `weights`) at `degree` passes, rtol 1e-8."""
    scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": degree,
             "linear_system/hsc/max_iter": 5000}
    ext = dict(C1_EXT, precision=precision, residual_replace_every=25, preconditioner=preconditioner)
    if preconditioner == "chebyshev":
        ext["poly_interval_ratio"] = float(ratio)
    h = plugin.Handle(plugin.make_config(scene, **ext))
    if weights is not None:
        w = torch.as_tensor(np.ascontiguousarray(weights, dtype=np.float32), device=device)
        h._check(h.lib.hsc_set_learned_weights(h.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
    A = plugin.BCOO3.from_system(sysm, device)
    h.setup(A, sysm.qo, sysm.qm, sysm.fo, sysm.fn)
    h._A = A
    return h


def sweep(sysm, degrees=DEGREES, ratios=RATIOS, device="cuda"):
    """PCG iterations of the Chebyshev-weighted solve (mixed precision, rtol
    1e-8) at every (degree, ratio); a solve that fails or does not converge
    counts as 10^4 iterations."""
    b = torch.as_tensor(sysm.b, device=device)
    x = torch.zeros_like(b)
    it = np.zeros((len(degrees), len(ratios)))
    for i, m in enumerate(degrees):
        for j, r in enumerate(ratios):
            h = poly_handle(sysm, m, r, device=device)
            st = h.solve(b, x, check=False)
            it[i, j] = st["pcg_iterations"] if st["status"] == 0 else 1e4
            h.close()
    return it


def best_log_ratio(iters, ratios=RATIOS) -> float:
    """log10 of the ratio with the fewest iterations: the middle of the
    tied minima, refined by the vertex of the parabola through the best
    point and its two neighbours (in log10 ratio), kept between them."""
    lr = np.log10(np.asarray(ratios))
    lo = np.flatnonzero(iters == iters.min())
    k = int(lo[len(lo) // 2])
    if len(lo) > 1 or k == 0 or k == len(lr) - 1:
        return float(np.mean(lr[lo]))
    x0, x1, x2 = lr[k - 1:k + 2]
    y0, y1, y2 = iters[k - 1:k + 2]
    den = (x0 - x1) * (x0 - x2) * (x1 - x2)
    a = (x2 * (y1 - y0) + x1 * (y0 - y2) + x0 * (y2 - y1)) / den
    bq = (x2 * x2 * (y0 - y1) + x1 * x1 * (y2 - y0) + x0 * x0 * (y1 - y2)) / den
    if a <= 0:
        return float(x1)
    return float(np.clip(-bq / (2 * a), x0, x2))


@dataclass
class TrainingSystem:
    key: tuple
    nb: int
    X: torch.Tensor        # [nb, 16] bf16 features
    ptr: torch.Tensor      # node graph
    nbr: torch.Tensor
    iters: np.ndarray      # [degrees, ratios] PCG iterations
    target: np.ndarray     # [3] channel targets log10(best ratio) - CENTER


def harvest(keys, device="cuda", log=None, workers=8) -> list:
    """Targets and features of the training systems (built on `workers` CPU
    processes, swept on the GPU)."""
    built = systems.map_processes(training_system, keys, workers)
    out = []
    for key, s in zip(keys, built):
        it = sweep(s, device=device)
        tgt = np.array([best_log_ratio(it[i]) - CENTER for i in range(len(DEGREES))])
        h = poly_handle(s, 16, device=device)
        ptr, nbr = node_graph(h, device)
        out.append(TrainingSystem(tuple(key), h.setup_info()["fem_nodes"], node_features(h, device), ptr, nbr, it,
                                  tgt))
        h.close()
        if log:
            log(f"{key}: nb {out[-1].nb}, best log10 ratio at degrees {DEGREES}: "
                f"{np.round(tgt + CENTER, 2).tolist()}, iterations {it.min(axis=1).astype(int).tolist()}")
    return out


def union_graph(data: list):
    """The disjoint union of the systems' node graphs, features and per-node
    targets (every node of a system carries that system's target)."""
    ptrs, nbrs, X, T = [], [], [], []
    node_off = edge_off = 0
    for d in data:
        ptrs.append(d.ptr[:-1].long() + edge_off)
        nbrs.append(d.nbr.long() + node_off)
        X.append(d.X)
        T.append(torch.as_tensor(d.target, dtype=torch.float32, device=d.X.device).expand(d.nb, 3))
        node_off += d.nb
        edge_off += int(d.nbr.numel())
    ptr = torch.cat(ptrs + [torch.tensor([edge_off], device=data[0].X.device)]).to(torch.int32)
    return ptr, torch.cat(nbrs).to(torch.int32), torch.cat(X)[None].contiguous(), torch.cat(T)[None].contiguous()


def train(data: list, steps=600, lr=3e-3, seed=0, log=None, theta0=None):
    """Full-batch training on the union graph with the library's BF16 kernels
    and Adam (cosine learning-rate schedule): per-node mean squared error of
    the three channels against the system targets.  Returns (theta, loss
    history)."""
    ptr, nbr, X, T = union_graph(data)
    net = GraphMLP(ptr, nbr, 1)
    net.set_params(init_params(seed) if theta0 is None else theta0)
    den = torch.tensor([float(X.shape[1])], dtype=torch.float32, device=X.device)
    hist = []
    t0 = time.perf_counter()
    for k in range(steps):
        lr_k = lr * 0.5 * (1 + np.cos(np.pi * k / steps))
        hist.append(net.train_step(X, T, den, lr_k))
        if log and (k % 100 == 0 or k == steps - 1):
            log(f"step {k}: loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    theta = net.get_params()
    net.close()
    return theta, hist


def predicted_log_ratios(theta, sysm, device="cuda") -> np.ndarray:
    """The network's three channel means + CENTER on a system (through a
    learned handle's setup)."""
    h = poly_handle(sysm, 16, preconditioner="learned", weights=theta, device=device)
    out = np.asarray(h.setup_info()["learned_out"]) + CENTER
    h.close()
    return out


def _median_solve(h, b, reps):
    x = torch.zeros_like(b)
    its, ms = [], []
    for _ in range(reps):
        st = h.solve(b, x, check=False)
        its.append(st["pcg_iterations"])
        ms.append(st["pcg_ms"])
    return {"status": st["status"], "iterations": int(np.median(its)), "pcg_ms": float(np.median(ms)),
            "schur_rel": st["schur_relative_residual"], "full_rel": st["full_relative_residual"],
            "fallback": st["fallback"], "learned_used": st["learned_used"]}


def evaluate(theta, sysms: dict, degree=16, device="cuda", reps=5):
    """PCG iterations and device ms (median of `reps` solves, mixed
    precision, rtol 1e-8) of the learned polynomial against the plain
    Neumann series (the method without the network: order 2 and omega 0.6
    for the 2,178-DOF family as configured for it, the drop-in defaults 5
    and 0.7 for the tori) and the Chebyshev passes at the fixed ratios 100
    (the config default) and 1000."""
    rows = []
    for name, s in sysms.items():
        neumann = (2, 0.6) if name.startswith(("cube", "c1")) else (5, 0.7)
        b = torch.as_tensor(s.b, device=device)
        out = {"system": name, "nb": s.fn // 3}
        scene = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": neumann[0],
                 "linear_system/hsc/omega": neumann[1], "linear_system/hsc/max_iter": 20000}
        hN = plugin.Handle(plugin.make_config(scene, precision="mixed", residual_replace_every=25, **C1_EXT))
        A = plugin.BCOO3.from_system(s, device)
        hN.setup(A, s.qo, s.qm, s.fo, s.fn)
        out["neumann"] = _median_solve(hN, b, reps)
        hN.close()
        for r in (100.0, 1000.0):
            h = poly_handle(s, degree, r, device=device)
            out[f"chebyshev_{int(r)}"] = _median_solve(h, b, reps)
            h.close()
        h = poly_handle(s, degree, preconditioner="learned", weights=theta, device=device)
        out["learned"] = _median_solve(h, b, reps)
        out["learned"]["ratio"] = h.setup_info()["poly_ratio"]
        h.close()
        rows.append(out)
    return rows


def save_weights(theta, path=WEIGHTS, meta=None):
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, theta=np.asarray(theta, dtype=np.float32), **(meta or {}))


def load_weights(path=WEIGHTS) -> np.ndarray:
    return np.load(path)["theta"]


def eval_systems(cube_seeds=range(200, 204), torus_seeds=range(200, 204), vehicle_seeds=range(200, 202),
                 vehicle=False) -> dict:
    """Held-out systems: seeds 200-203 of the cube and torus families, 200-201
    of the tire family, the 2,178-DOF system itself and (optionally) the
    78k-tet system."""
    out = {f"cube{sd}": family_system(sd) for sd in cube_seeds}
    out["c1"] = systems.cube_system(seed=0)
    out.update({f"torus{sd}": torus_family_system(sd) for sd in torus_seeds})
    out.update({f"vehicle{sd}": vehicle_family_system(sd) for sd in vehicle_seeds})
    if vehicle:
        out["c2"] = systems.torus_system(seed=1)
    return out


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="sweep, train and evaluate the learned polynomial preconditioner")
    ap.add_argument("--cube-seeds", type=int, nargs=2, default=[100, 200], help="training seeds [a, b) of the cubes")
    ap.add_argument("--torus-seeds", type=int, nargs=2, default=[100, 150], help="training seeds [a, b) of the tori")
    ap.add_argument("--vehicle-seeds", type=int, nargs=2, default=[100, 130],
                    help="training seeds [a, b) of the tire systems at the 78k-tet system's resolution")
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--report", default=None, help="JSON file for the sweep targets and the evaluation")
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    keys = ([("cube", sd) for sd in range(*a.cube_seeds)] + [("torus", sd) for sd in range(*a.torus_seeds)]
            + [("vehicle", sd) for sd in range(*a.vehicle_seeds)])
    t0 = time.perf_counter()
    data = harvest(keys, log=log, workers=a.workers)
    log(f"swept {len(data)} systems ({sum(d.nb for d in data)} nodes) in {time.perf_counter() - t0:.1f} s")
    theta, hist = train(data, steps=a.steps, lr=a.lr, log=log)
    evs = eval_systems(vehicle=True)
    rows = evaluate(theta, evs)
    for r in rows:
        log(json.dumps(r))
    save_weights(theta, Path(a.out), meta={"cube_seeds": np.array(a.cube_seeds), "torus_seeds": np.array(a.torus_seeds),
                                           "vehicle_seeds": np.array(a.vehicle_seeds), "steps": np.array(a.steps),
                                           "loss": np.array(hist)})
    if a.report:
        Path(a.report).write_text(json.dumps({
            "targets": [{"key": list(d.key), "nb": d.nb, "target_log10_ratio": (d.target + CENTER).tolist(),
                         "min_iterations": d.iters.min(axis=1).tolist()} for d in data],
            "loss_first_last": [hist[0], hist[-1]], "evaluation": rows}, indent=1))
    log(f"saved {a.out}")


if __name__ == "__main__":
    main()
