"""Host-side mirror of the reference's linear-solver plugin for the HSC path.

The reference plugs ``LinearHSC`` into libuipc's ``IterativeSolver`` slot
(reproduction/hsc_backend/linear_hsc.cu:153-438;
upstream/src/backends/cuda/linear_system/iterative_solver.h:19-21):

* ``do_build`` reads ``linear_system/solver`` (must be "linear_hsc", else
  the system is dropped) and the ``linear_system/hsc/*`` keys
  (linear_hsc.cu:172-193; defaults in integration.patch:34-55);
* ``do_solve(SolvingInfo)`` solves H x = b for the Newton direction given
  ``info.A()`` (upper BCOO of 3x3 blocks), ``info.b()`` and writes
  ``info.x()``, ``info.iter_count()``, and appends one JSONL row to
  ``workspace()/hsc_stats.jsonl`` (linear_hsc.cu:344-436).

Here the same contract is expressed over the C ABI (include/hsc.h) with
torch device tensors as the buffers.  Errors raise ``HSCError`` carrying the
status code -- the reference's UIPC_ASSERT failures, one code each.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass
from pathlib import Path

import torch

from . import _abi

# Library objects still alive at interpreter exit are destroyed by an atexit
# hook (registered after torch's, so it runs first), while the HIP runtime is
# certainly up: destructors left to module teardown could otherwise call HIP
# after a profiler or the runtime has finalised.  Batches go before the
# handles they hold, then networks, then handles.
_LIVE = {"batch": weakref.WeakSet(), "mlp": weakref.WeakSet(), "handle": weakref.WeakSet()}


def register_live(kind: str, obj) -> None:
    _LIVE[kind].add(obj)


@atexit.register
def _close_live() -> None:
    for kind in ("batch", "mlp", "handle"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass

# libuipc scene-config keys of the plugin and their defaults
CONFIG_DEFAULTS = {
    "linear_system/solver": "linear_hsc",
    "linear_system/tol_rate": 1e-3,
    "linear_system/hsc/max_rank": 128,
    "linear_system/hsc/neumann_order": 5,
    "linear_system/hsc/max_iter": 2000,
    "linear_system/hsc/adaptive_abd_inverse": 1,
    "linear_system/hsc/abd_inverse_tol": 1e-5,
    "linear_system/hsc/full_residual_limit": 0.1,
    "linear_system/hsc/omega": 0.7,
}


class SimSystemException(RuntimeError):
    """Raised by do_build when the plugin is not selected (linear_hsc.cu:176)."""


class HSCError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{_abi.STATUS.get(code, code)}: {message}")
        self.code = code
        self.name = _abi.STATUS.get(code, str(code))


def make_config(scene: dict | None = None, **extra) -> _abi.Config:
    """Build an hsc_config from libuipc-style keys plus extension fields
    (abd_inverse, precision, residual_replace_every,
    true_residual_every_iteration, aca_rtol, lanes_per_row, wg_per_cu, schedule,
    poly_interval_ratio, pass_sync, pipelined, operator_mapping,
    preconditioner, check_acceptance)."""
    lib = _abi.load()
    cfg = _abi.Config()
    lib.hsc_config_default(C.byref(cfg))
    sc = dict(CONFIG_DEFAULTS)
    sc.update(scene or {})
    cfg.tol_rate = float(sc["linear_system/tol_rate"])
    cfg.max_rank = int(sc["linear_system/hsc/max_rank"])
    cfg.neumann_order = int(sc["linear_system/hsc/neumann_order"])
    cfg.max_iter = int(sc["linear_system/hsc/max_iter"])
    cfg.adaptive_abd_inverse = int(sc["linear_system/hsc/adaptive_abd_inverse"])
    cfg.abd_inverse_tol = float(sc["linear_system/hsc/abd_inverse_tol"])
    cfg.full_residual_limit = float(sc["linear_system/hsc/full_residual_limit"])
    cfg.omega = float(sc["linear_system/hsc/omega"])
    for key, val in extra.items():
        if key == "precision":
            val = _abi.PRECISION[val] if isinstance(val, str) else val
        elif key == "abd_inverse":
            val = _abi.ABD_INVERSE[val] if isinstance(val, str) else val
        elif key == "preconditioner":
            val = _abi.PRECONDITIONER[val] if isinstance(val, str) else val
        if not hasattr(cfg, key):
            raise KeyError(key)
        setattr(cfg, key, val)
    return cfg


def _ptr(t: torch.Tensor) -> int:
    if t.device.type != "cuda":
        raise ValueError("HSC buffers must be device tensors")
    if not t.is_contiguous():
        raise ValueError("HSC buffers must be contiguous")
    return t.data_ptr()


def _stream(stream) -> C.c_void_p:
    if stream is None:
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


@dataclass
class BCOO3:
    """Device upper-triangle BCOO of 3x3 blocks (libuipc CBCOOMatrixView<double,3>)."""

    rows: torch.Tensor    # int32 [nnz]
    cols: torch.Tensor    # int32 [nnz]
    values: torch.Tensor  # float64 [nnz, 9] column-major blocks
    block_dim: int

    def view(self) -> _abi.BCOO3View:
        n = int(self.rows.numel())
        return _abi.BCOO3View(_ptr(self.rows) if n else 0, _ptr(self.cols) if n else 0,
                              _ptr(self.values) if n else 0, n, int(self.block_dim))

    @classmethod
    def from_system(cls, sysm, device="cuda") -> "BCOO3":
        return cls(torch.as_tensor(sysm.rows, dtype=torch.int32, device=device).contiguous(),
                   torch.as_tensor(sysm.cols, dtype=torch.int32, device=device).contiguous(),
                   torch.as_tensor(sysm.vals, dtype=torch.float64, device=device).contiguous(),
                   sysm.N // 3)


def fem_hessian(tets: torch.Tensor, x: torch.Tensor, Dm_inv: torch.Tensor, mu: torch.Tensor, lam: torch.Tensor,
                volumes: torch.Tensor, dt: float = 1.0, project_spd: bool = True, stream=None):
    """Per-tet log-Neo-Hookean Hessian blocks on the device (hsc_fem_hessian):
    returns (rows, cols, values (10 n, 9) column-major), 10 general triplets
    per tet for assemble_bcoo.  Dm_inv: (n, 3, 3) row-major tensors (stored
    column-major for the kernel)."""
    lib = _abi.load()
    n = int(tets.shape[0])
    dev = tets.device
    t = tets.to(torch.int32).contiguous()
    xx = x.to(torch.float64).contiguous()
    dm = Dm_inv.to(torch.float64).transpose(1, 2).contiguous()  # column-major 3x3 per tet
    m = mu.to(torch.float64).contiguous()
    la = lam.to(torch.float64).contiguous()
    vo = volumes.to(torch.float64).contiguous()
    rows = torch.empty(max(10 * n, 1), dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty((max(10 * n, 1), 9), dtype=torch.float64, device=dev)
    if n:
        code = lib.hsc_fem_hessian(_ptr(t), _ptr(xx), _ptr(dm), _ptr(m), _ptr(la), _ptr(vo), float(dt), n,
                                   1 if project_spd else 0, _ptr(rows), _ptr(cols), _ptr(vals), _stream(stream))
        if code != _abi.HSC_OK:
            raise HSCError(code, "hsc_fem_hessian failed")
    return rows[:10 * n], cols[:10 * n], vals[:10 * n]


def abd_fem_coupling(abd_vertex: torch.Tensor, fem_vertex: torch.Tensor, hess: torch.Tensor,
                     v2b: torch.Tensor, xbar: torch.Tensor, body_fixed: torch.Tensor, vertex_fixed: torch.Tensor,
                     abd_block_offset: int, fem_block_offset: int, stream=None):
    """ABD-FEM contact blocks -> the coupling triplets of the plugin's input on
    the device (hsc_abd_fem_coupling; abd_fem_linear_subsystem.cu:18-61):
    returns (rows, cols, values (4 n, 9) column-major), 4 triplets per contact
    block for assemble_bcoo.  hess: (n, 3, 3) row-major contact blocks (ABD
    vertex rows, FEM vertex columns); xbar: (abd vertices, 3) rest positions;
    body_fixed / vertex_fixed: per body / per FEM vertex flags."""
    lib = _abi.load()
    n = int(abd_vertex.numel())
    dev = abd_vertex.device
    av = abd_vertex.to(torch.int32).contiguous()
    fv = fem_vertex.to(torch.int32).contiguous()
    h = hess.to(torch.float64).reshape(n, 3, 3).transpose(1, 2).contiguous()  # column-major 3x3
    vb = v2b.to(torch.int32).contiguous()
    xb = xbar.to(torch.float64).reshape(-1, 3).contiguous()
    bf = body_fixed.to(torch.int32).contiguous()
    vf = vertex_fixed.to(torch.int32).contiguous()
    rows = torch.empty(max(4 * n, 1), dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty((max(4 * n, 1), 9), dtype=torch.float64, device=dev)
    if n:
        code = lib.hsc_abd_fem_coupling(_ptr(av), _ptr(fv), _ptr(h), n, _ptr(vb), _ptr(xb), int(vb.numel()),
                                        _ptr(bf), int(bf.numel()), _ptr(vf), int(vf.numel()),
                                        int(abd_block_offset), int(fem_block_offset), _ptr(rows), _ptr(cols),
                                        _ptr(vals), _stream(stream))
        if code != _abi.HSC_OK:
            raise HSCError(code, "hsc_abd_fem_coupling failed")
    return rows[:4 * n], cols[:4 * n], vals[:4 * n]


def assemble_bcoo(rows: torch.Tensor, cols: torch.Tensor, values: torch.Tensor, block_dim: int,
                  fold_lower: bool = True, stream=None) -> BCOO3:
    """General 3x3-block triplets on the device (duplicates allowed) -> the
    upper-triangle BCOO of unique blocks the plugin takes (hsc_assemble_bcoo;
    the caller-side ge2sym + convert of global_linear_system.cu:232-234)."""
    lib = _abi.load()
    n = int(rows.numel())
    dev = rows.device
    r = rows.to(torch.int32).contiguous()
    c = cols.to(torch.int32).contiguous()
    v = values.to(torch.float64).reshape(n, 9).contiguous()
    orow = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ocol = torch.empty_like(orow)
    oval = torch.empty((max(n, 1), 9), dtype=torch.float64, device=dev)
    cnt = C.c_int64()
    code = lib.hsc_assemble_bcoo(_ptr(r) if n else None, _ptr(c) if n else None, _ptr(v) if n else None, n,
                                 1 if fold_lower else 0, _ptr(orow), _ptr(ocol), _ptr(oval), C.byref(cnt),
                                 _stream(stream))
    if code != _abi.HSC_OK:
        raise HSCError(code, "hsc_assemble_bcoo failed")
    k = cnt.value
    return BCOO3(orow[:k].clone(), ocol[:k].clone(), oval[:k].clone(), block_dim)


@dataclass
class SolvingInfo:
    """Mirror of GlobalLinearSystem::SolvingInfo (global_linear_system.h:227-245
    plus A(), integration.patch:7): x (out), b, A, and the DOF ranges the
    reference reads from its ABD/FEM subsystems (linear_hsc.cu:197-198)."""

    x: torch.Tensor
    b: torch.Tensor
    A: BCOO3
    qo: int
    qm: int
    fo: int
    fn: int
    iter_count: int = 0
    frame: int = 0
    newton: int = 0
    # contact counters of the JSONL row: the reference reads them from its
    # contact systems (linear_hsc.cu:410-414); the caller supplies them here
    active_mesh_contacts: int = 0
    active_ground_contacts: int = 0


class Handle:
    """Owns one hsc_handle (one system at a time, one stream at a time)."""

    def __init__(self, config: _abi.Config | None = None, device: int | None = None):
        self.lib = _abi.load()
        self.device = torch.cuda.current_device() if device is None else int(device)
        self.config = config if config is not None else make_config()
        h = C.c_void_p()
        code = self.lib.hsc_create(C.byref(self.config), self.device, C.byref(h))
        if code != 0:
            raise HSCError(code, "hsc_create")
        self.h = h
        register_live("handle", self)

    def _check(self, code: int) -> None:
        if code != 0:
            msg = self.lib.hsc_last_error(self.h)
            raise HSCError(code, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.hsc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, cfg: _abi.Config):
        self._check(self.lib.hsc_set_config(self.h, C.byref(cfg)))
        self.config = cfg

    def setup(self, A: BCOO3, qo, qm, fo, fn, stream=None):
        v = A.view()
        self._keep = A
        self._check(self.lib.hsc_setup(self.h, C.byref(v), qo, qm, fo, fn, _stream(stream)))

    def solve(self, b: torch.Tensor, x: torch.Tensor, stream=None, check=True) -> dict:
        st = _abi.Stats()
        code = self.lib.hsc_solve(self.h, C.c_void_p(_ptr(b)), C.c_void_p(_ptr(x)), C.byref(st), _stream(stream))
        if check:
            self._check(code)
        d = st.as_dict()
        d["status_name"] = _abi.STATUS.get(code, str(code))
        return d

    def apply_D(self, x, y, stream=None):
        self._check(self.lib.hsc_apply_D(self.h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y)), _stream(stream)))

    def apply_schur(self, x, y, stream=None):
        self._check(self.lib.hsc_apply_schur(self.h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y)), _stream(stream)))

    def apply_precond(self, r, z, stream=None):
        self._check(self.lib.hsc_apply_precond(self.h, C.c_void_p(_ptr(r)), C.c_void_p(_ptr(z)), _stream(stream)))

    def apply_pass(self, x, r, y, w: float, stream=None):
        """y = x + w Dinv (r - D x) (hsc_apply_pass): x, y float64 with the FP64
        values, or float32 with the FP32 copies; r float64."""
        import torch
        if x.dtype != y.dtype or x.dtype not in (torch.float64, torch.float32) or r.dtype != torch.float64:
            raise TypeError("x and y: both float64 or both float32; r: float64")
        prec = _abi.PRECISION["fp64"] if x.dtype == torch.float64 else _abi.PRECISION["mixed"]
        self._check(self.lib.hsc_apply_pass(self.h, prec, float(w), C.c_void_p(_ptr(x)), C.c_void_p(_ptr(r)),
                                            C.c_void_p(_ptr(y)), _stream(stream)))

    def setup_info(self) -> dict:
        si = _abi.SetupInfo()
        self._check(self.lib.hsc_get_setup_info(self.h, C.byref(si)))
        return si.as_dict()

    TIMED_KERNELS = {"spmv64_rows": 0, "pass64_rows": 1, "pass32_rows": 2, "spmv64": 3, "pass64": 4, "pass32": 5}

    def time_kernel(self, kernel: str, shape: int = 0, repeats: int = 50, stream=None) -> float:
        """ms per launch of one standalone operator kernel on this handle's D
        (hsc_time_kernel): "spmv64" (y = D x), "pass64" / "pass32" (one
        block-Jacobi pass in FP64 / FP32) with the streaming kernels, `shape`
        64-block chunks per wavefront (0: the library's choice); the "_rows"
        forms map `shape` (4, 8, 16) lanes to each block row."""
        ms = C.c_double()
        self._check(self.lib.hsc_time_kernel(self.h, self.TIMED_KERNELS[kernel], shape, repeats,
                                             _stream(stream), C.byref(ms)))
        return ms.value

    def factors(self) -> dict:
        import numpy as np
        info = self.setup_info()
        qm, a, k = info["abd_dofs"], info["active_fem_dofs"], info["rank"]
        U = np.zeros((k, qm))          # column-major qm x k == row-major k x qm
        V = np.zeros((k, max(a, 0)))
        Cm = np.zeros((k, k))
        P = np.zeros((qm, qm))
        act = np.zeros(max(a, 0), dtype=np.int32)
        hist = np.zeros(k + 1)
        self._check(self.lib.hsc_get_factors(
            self.h, U.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p), Cm.ctypes.data_as(C.c_void_p),
            P.ctypes.data_as(C.c_void_p), act.ctypes.data_as(C.c_void_p), hist.ctypes.data_as(C.c_void_p)))
        return {"U": U.T.copy(), "V": V.T.copy(), "C": Cm, "P": P, "active": act, "aca_history": hist, **info}


def solve_batched(handles: list[Handle], bs: list[torch.Tensor], xs: list[torch.Tensor], stream=None) -> list[dict]:
    """Solve many set-up systems in ONE persistent launch."""
    lib = _abi.load()
    n = len(handles)
    hs = (C.c_void_p * n)(*[h.h.value for h in handles])
    bp = (C.c_void_p * n)(*[_ptr(b) for b in bs])
    xp = (C.c_void_p * n)(*[_ptr(x) for x in xs])
    stats = (_abi.Stats * n)()
    code = lib.hsc_solve_batched(hs, n, bp, xp, stats, _stream(stream))
    out = []
    for i in range(n):
        d = stats[i].as_dict()
        d["status_name"] = _abi.STATUS.get(d["status"], str(d["status"]))
        out.append(d)
    if code != 0 and all(d["status"] == 0 for d in out):
        msg = lib.hsc_last_error(handles[0].h)
        raise HSCError(code, msg.decode() if msg else "")
    return out


class Batch:
    """A batched solve captured once into a hipGraph (hsc_batch_*), replayed
    for the same b/x buffers.  Keeps its handles and buffers alive."""

    def __init__(self, handles: list[Handle], bs: list[torch.Tensor], xs: list[torch.Tensor], stream=None):
        self.lib = _abi.load()
        self.handles, self.bs, self.xs = list(handles), list(bs), list(xs)
        n = len(handles)
        hs = (C.c_void_p * n)(*[h.h.value for h in handles])
        bp = (C.c_void_p * n)(*[_ptr(b) for b in bs])
        xp = (C.c_void_p * n)(*[_ptr(x) for x in xs])
        out = C.c_void_p()
        code = self.lib.hsc_batch_create(hs, n, bp, xp, _stream(stream), C.byref(out))
        if code != 0:
            msg = self.lib.hsc_last_error(handles[0].h)
            raise HSCError(code, msg.decode() if msg else "hsc_batch_create")
        self.b = out
        register_live("batch", self)

    def launch(self, stream=None):
        code = self.lib.hsc_batch_launch(self.b, _stream(stream))
        if code != 0:
            msg = self.lib.hsc_last_error(self.handles[0].h)
            raise HSCError(code, msg.decode() if msg else "hsc_batch_launch")

    def stats(self, stream=None, check=True) -> list[dict]:
        n = len(self.handles)
        st = (_abi.Stats * n)()
        code = self.lib.hsc_batch_stats(self.b, st, _stream(stream))
        out = []
        for i in range(n):
            d = st[i].as_dict()
            d["status_name"] = _abi.STATUS.get(d["status"], str(d["status"]))
            out.append(d)
        if check and code != 0:
            msg = self.lib.hsc_last_error(self.handles[0].h)
            raise HSCError(code, msg.decode() if msg else "hsc_batch_stats")
        return out

    def info(self) -> dict:
        v = [C.c_int32() for _ in range(5)]
        self.lib.hsc_batch_info(self.b, *[C.byref(x) for x in v])
        return dict(zip(("lanes_per_row", "wg_per_cu", "workgroups", "cooperative_node", "schedule"),
                        [x.value for x in v]))

    def close(self):
        if getattr(self, "b", None):
            self.lib.hsc_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LinearHSC:
    """Plugin-shaped adapter: do_build(scene config) / do_solve(SolvingInfo)."""

    def __init__(self, workspace: str | Path | None = None, **extensions):
        self.workspace = Path(workspace) if workspace else None
        self.extensions = extensions
        self.handle: Handle | None = None

    def do_build(self, scene_config: dict) -> None:
        if scene_config.get("linear_system/solver", "linear_hsc") != "linear_hsc":
            raise SimSystemException("LinearHSC unused")
        cfg = make_config(scene_config, **self.extensions)
        self.handle = Handle(cfg)

    def do_solve(self, info: SolvingInfo) -> dict:
        if self.handle is None:
            raise RuntimeError("do_build first")
        lib = self.handle.lib
        st = _abi.Stats()
        v = info.A.view()
        code = lib.hsc_do_solve(self.handle.h, C.byref(v), info.qo, info.qm, info.fo, info.fn,
                                C.c_void_p(_ptr(info.b)), C.c_void_p(_ptr(info.x)), C.byref(st), _stream(None))
        info.iter_count = st.pcg_iterations
        st.active_mesh_contacts = int(info.active_mesh_contacts)
        st.active_ground_contacts = int(info.active_ground_contacts)
        if self.workspace is not None:
            self.workspace.mkdir(parents=True, exist_ok=True)
            lib.hsc_append_jsonl(C.byref(st), str(self.workspace / "hsc_stats.jsonl").encode(),
                                 info.frame, info.newton)
        self.handle._check(code)
        return st.as_dict()
