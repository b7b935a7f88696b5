"""Learned correction of the Neumann preconditioner (no reference counterpart).

A per-node graph MLP over the FEM node graph of D (two message-passing layers,
hidden width 64, BF16 weights and activations, FP32 accumulation) maps the
local diagonal block and residual to a 3-vector correction of the Neumann
preconditioner's output:  z_i = zN_i + (rms(r) / s_i) o_i.  It runs inside
flexible PCG (Polak-Ribiere beta); when the corrected z is not a descent
direction the plain Neumann z is used for that iteration (the fallback).

Everything numeric runs in the HIP library through the C ABI:
  * harvesting: hsc_solve_traced stores (r_k, x_k) of the plain solver's own
    PCG iterations; e_k = x* - x_k;
  * training signal: hsc_energy_targets (energy norm |e - z|_S^2, S applied
    by the solver's own Schur operator);
  * training: hsc_mlp_train_step (hand-written BF16 forward/backward + Adam);
  * inference: hsc_set_learned_weights + preconditioner = "learned".
torch tensors are only the device buffers.

Train the committed weights with ``python -m hsc_reimpl_amd.learned``.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _abi, plugin, systems

WEIGHTS = Path(__file__).resolve().parent / "data" / "learned_graph_mlp.npz"
C1_SCENE = {"linear_system/tol_rate": 1e-8, "linear_system/hsc/neumann_order": 2,
            "linear_system/hsc/omega": 0.6, "linear_system/hsc/max_iter": 5000}
C1_EXT = {"aca_rtol": 1e-6}


def param_count() -> int:
    return int(_abi.load().hsc_mlp_param_count())


def init_params(seed=0) -> np.ndarray:
    """Flat float32 parameters in the library's layout (He init; output
    layer scaled by 0.1; padded rows/columns zero)."""
    rng = np.random.default_rng(seed)
    HID, IN, OUTP = 64, 16, 32
    parts = []
    W0 = rng.normal(size=(IN, HID)) * np.sqrt(2.0 / 12)
    W0[12:] = 0.0
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


def node_features(handle: plugin.Handle, device="cuda"):
    info = handle.setup_info()
    nb = info["fem_nodes"]
    dfeat = torch.zeros((nb, 9), dtype=torch.float32, device=device)
    sdiag = torch.zeros(nb, dtype=torch.float32, device=device)
    handle._check(handle.lib.hsc_get_node_features(handle.h, C.c_void_p(dfeat.data_ptr()),
                                                   C.c_void_p(sdiag.data_ptr())))
    return dfeat, sdiag


def features(dfeat: torch.Tensor, r: torch.Tensor):
    """X [B, nb, 16] bf16 and rho [B] from D_ii/s_i and residuals r [B, 3nb]."""
    B = r.shape[0]
    nb = dfeat.shape[0]
    rr = r.view(B, nb, 3)
    rho = torch.sqrt(torch.mean(rr * rr, dim=(1, 2)))
    rho = torch.where(rho > 0, rho, torch.ones_like(rho))
    X = torch.zeros((B, nb, 16), dtype=torch.float32, device=r.device)
    X[:, :, :9] = dfeat[None]
    X[:, :, 9:12] = (rr / rho[:, None, None]).float()
    return X.to(torch.bfloat16), rho


@dataclass
class SystemPairs:
    """Training pairs of one system of the family (kept with its handle,
    whose Schur operator defines the energy norm)."""
    seed: int
    handle: plugin.Handle
    A: plugin.BCOO3
    X: torch.Tensor      # [P, nb, 16] bf16 features of r
    U: torch.Tensor      # [P, n] fp64: e - Neumann(r), the part of the error the network should add
    rho: np.ndarray      # [P] rms(r)
    den: np.ndarray      # [P] U^T S U (energy of the Neumann error)

    def __len__(self):
        return self.X.shape[0]


def family_system(seed: int):
    """A system of the 2,178-DOF family with E log-uniform in [1e6, 1e10]."""
    E = float(np.exp(np.random.default_rng(10_000 + seed).uniform(np.log(1e6), np.log(1e10))))
    return systems.cube_system(seed=seed, E=E)


def energy_targets(sp: SystemPairs, idx, O=None, T=None) -> np.ndarray:
    """Per-sample energies |U - (rho/s) o|_S^2 (and T for hsc_mlp_train_step)."""
    idx = np.asarray(idx)
    U = sp.U[torch.as_tensor(idx, device=sp.U.device)].contiguous()
    rho = np.ascontiguousarray(sp.rho[idx], dtype=np.float64)
    den = np.ascontiguousarray(sp.den[idx], dtype=np.float64) if T is not None else None
    en = np.zeros(len(idx))
    h = sp.handle
    h._check(h.lib.hsc_energy_targets(
        h.h, C.c_void_p(O.data_ptr()) if O is not None else None, C.c_void_p(U.data_ptr()),
        rho.ctypes.data_as(C.c_void_p), den.ctypes.data_as(C.c_void_p) if den is not None else None, len(idx),
        C.c_void_p(T.data_ptr()) if T is not None else None, en.ctypes.data_as(C.c_void_p), plugin._stream(None)))
    return en


def harvest(seeds, per_seed=200, device="cuda", log=None, max_rhs=4) -> list:
    """Residual/error pairs (r_k, e_k = x* - x_k) from the plain solver's own
    traced PCG iterations, on the system's b and further N(0,1) right-hand
    sides until `per_seed` pairs are collected.  Pairs whose residual fell
    below 1e-6 of the initial one are dropped (their error is dominated by
    the solve's own tolerance)."""
    lib = _abi.load()
    out = []
    for seed in seeds:
        s = family_system(seed)
        h = plugin.Handle(plugin.make_config(C1_SCENE, **C1_EXT))
        A = plugin.BCOO3.from_system(s, device)
        h.setup(A, s.qo, s.qm, s.fo, s.fn)
        n = s.fn
        rng = np.random.default_rng(20_000 + seed)
        Rs, Es = [], []
        got = 0
        for j in range(max_rhs):
            b = torch.as_tensor(s.b if j == 0 else rng.normal(size=s.N), device=device)
            x = torch.zeros_like(b)
            slots = 4 * per_seed
            tr = torch.zeros((slots, n), dtype=torch.float64, device=device)
            tx = torch.zeros((slots, n), dtype=torch.float64, device=device)
            ns = C.c_int32()
            st = _abi.Stats()
            h._check(lib.hsc_solve_traced(h.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                          C.c_void_p(tr.data_ptr()), C.c_void_p(tx.data_ptr()), 1, slots,
                                          C.byref(ns), C.byref(st), plugin._stream(None)))
            k = ns.value
            R = tr[:k]
            Ev = x[s.fo:s.fo + n][None, :] - tx[:k]
            keep = torch.linalg.vector_norm(R, dim=1) >= 1e-6 * torch.linalg.vector_norm(R[0])
            Rs.append(R[keep])
            Es.append(Ev[keep])
            got += int(keep.sum())
            if got >= per_seed:
                break
        R = torch.cat(Rs)[:per_seed]
        Ev = torch.cat(Es)[:per_seed]
        dfeat, _ = node_features(h, device)
        zN = torch.zeros_like(R)
        for i in range(R.shape[0]):
            h.apply_precond(R[i].contiguous(), zN[i])
        X, rho = features(dfeat, R)
        sp = SystemPairs(seed, h, A, X, (Ev - zN).contiguous(), rho.cpu().numpy(), np.zeros(R.shape[0]))
        sp.den = energy_targets(sp, np.arange(len(sp)))
        out.append(sp)
        if log:
            log(f"harvest seed {seed}: {len(sp)} pairs")
    return out


def train(data: list, graph, epochs=20, batch=32, lr=2e-3, seed=0, log=None, theta0=None):
    """Train on the GPU with the library's BF16 kernels, minimising the mean
    of |e - z|_S^2 / |e - Neumann(r)|_S^2 over pairs (z = Neumann(r) plus the
    network's correction), with Adam and a cosine learning-rate schedule.
    Batches never mix systems.  Returns (theta, per-epoch mean loss)."""
    ptr, nbr = graph
    net = GraphMLP(ptr, nbr, batch)
    net.set_params(init_params(seed) if theta0 is None else theta0)
    rng = np.random.default_rng(seed)
    batches = [(j, np.arange(i, min(i + batch, len(sp)))) for j, sp in enumerate(data)
               for i in range(0, len(sp), batch)]
    ones = torch.ones(batch, dtype=torch.float32, device=data[0].X.device)
    hist = []
    for ep in range(epochs):
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))
        tot = 0.0
        t0 = time.perf_counter()
        for bi in rng.permutation(len(batches)):
            j, idx = batches[bi]
            sp = data[j]
            X = sp.X[torch.as_tensor(idx, device=sp.X.device)].contiguous()
            O = net.forward(X)
            T = torch.empty_like(O)
            en = energy_targets(sp, idx, O, T)
            tot += float(np.mean(en / sp.den[idx]))
            net.train_step(X, T, ones[:len(idx)], lr_ep)
        hist.append(tot / len(batches))
        if log:
            log(f"epoch {ep}: energy loss {hist[-1]:.4f} ({time.perf_counter() - t0:.2f} s)")
    return net.get_params(), hist


def evaluate(theta, seeds, device="cuda"):
    """Iterations and time of flexible PCG with the learned correction vs the
    same host-driven PCG without it (identical operators, rtol 1e-8)."""
    lib = _abi.load()
    rows = []
    w = torch.as_tensor(np.ascontiguousarray(theta, dtype=np.float32), device=device)
    for seed in seeds:
        s = family_system(seed) if seed != "c1" else systems.cube_system(seed=0)
        A = plugin.BCOO3.from_system(s, device)
        b = torch.as_tensor(s.b, device=device)
        out = {"seed": seed}
        for mode in ("plain", "learned"):
            ext = dict(C1_EXT, preconditioner="learned") if mode == "learned" else dict(C1_EXT)
            h = plugin.Handle(plugin.make_config(C1_SCENE, **ext))
            h.setup(A, s.qo, s.qm, s.fo, s.fn)
            x = torch.zeros_like(b)
            st = _abi.Stats()
            if mode == "learned":
                h._check(lib.hsc_set_learned_weights(h.h, C.c_void_p(w.data_ptr()), w.numel() * 4))
                t0 = time.perf_counter()
                d = h.solve(b, x, check=False)
            else:
                t0 = time.perf_counter()
                code = lib.hsc_solve_traced(h.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), None, None, 0,
                                            0, None, C.byref(st), plugin._stream(None))
                d = st.as_dict()
                d["status"] = code
            el = time.perf_counter() - t0
            out[mode] = {"iterations": d["pcg_iterations"], "ms": 1e3 * el, "status": d["status"],
                         "schur_rel": d["schur_relative_residual"], "fallback": d["fallback"],
                         "full_rel": d["full_relative_residual"]}
        rows.append(out)
    return rows


def save_weights(theta, path=WEIGHTS, meta=None):
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, theta=np.asarray(theta, dtype=np.float32), **(meta or {}))


def load_weights(path=WEIGHTS) -> np.ndarray:
    return np.load(path)["theta"]


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="harvest, train and evaluate the learned preconditioner")
    ap.add_argument("--seeds", type=int, nargs=2, default=[100, 200], help="training seed range [a, b)")
    ap.add_argument("--per-seed", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-seeds", type=int, nargs=2, default=[200, 204])
    ap.add_argument("--out", default=str(WEIGHTS))
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    t0 = time.perf_counter()
    data = harvest(range(*a.seeds), per_seed=a.per_seed, log=log)
    npairs = sum(len(sp) for sp in data)
    log(f"harvested {npairs} pairs in {time.perf_counter() - t0:.1f} s")
    graph = node_graph(data[0].handle)
    t0 = time.perf_counter()
    theta, hist = train(data, graph, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    log(f"trained in {time.perf_counter() - t0:.1f} s")
    rows = evaluate(theta, list(range(*a.eval_seeds)) + ["c1"])
    for r in rows:
        log(json.dumps(r))
    save_weights(theta, Path(a.out), meta={"seeds": np.array(a.seeds), "pairs": np.array(npairs),
                                           "epochs": np.array(a.epochs), "loss": np.array(hist),
                                           "evaluation": np.array(json.dumps(rows))})
    log(f"saved {a.out}")


if __name__ == "__main__":
    main()
