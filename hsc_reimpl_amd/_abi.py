"""ctypes binding of the C ABI in include/hsc.h (libhsc_mi355x.so).

There is no fallback: if the HIP library is missing, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhsc_mi355x.so"

HSC_ABI_VERSION = 3

STATUS = {
    0: "HSC_OK", 1: "HSC_ERR_INVALID_ARGUMENT", 2: "HSC_ERR_FEM_DIAGONAL",
    3: "HSC_ERR_UNEXPECTED_COUPLING", 4: "HSC_ERR_ABD_NOT_SPD", 5: "HSC_ERR_ABD_INVERSE",
    6: "HSC_ERR_NONFINITE_FACTORS", 7: "HSC_ERR_BREAKDOWN", 8: "HSC_ERR_NOT_CONVERGED",
    9: "HSC_ERR_NOT_DESCENT", 10: "HSC_ERR_FULL_RESIDUAL", 11: "HSC_ERR_HIP",
    12: "HSC_ERR_NOT_SETUP", 13: "HSC_ERR_DEVICE_TIMEOUT", 14: "HSC_ERR_CAPACITY",
    15: "HSC_ERR_STALE_BATCH",
}
globals().update({v: k for k, v in STATUS.items()})

PRECISION = {"fp64": 0, "mixed": 1}
ABD_INVERSE = {"exact": 0, "neumann": 1, "neumann_adaptive": 2}
PRECONDITIONER = {"neumann": 0, "learned": 1, "chebyshev": 2}


class Config(C.Structure):
    _fields_ = [
        ("max_rank", C.c_int32), ("neumann_order", C.c_int32), ("max_iter", C.c_int32),
        ("adaptive_abd_inverse", C.c_int32), ("abd_inverse_tol", C.c_double),
        ("full_residual_limit", C.c_double), ("omega", C.c_double), ("tol_rate", C.c_double),
        ("abd_inverse", C.c_int32), ("precision", C.c_int32), ("residual_replace_every", C.c_int32),
        ("true_residual_every_iteration", C.c_int32), ("aca_rtol", C.c_double),
        ("lanes_per_row", C.c_int32), ("wg_per_cu", C.c_int32), ("preconditioner", C.c_int32),
        ("check_acceptance", C.c_int32), ("schedule", C.c_int32), ("poly_interval_ratio", C.c_double),
        ("pass_sync", C.c_int32), ("pipelined", C.c_int32), ("operator_mapping", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class BCOO3View(C.Structure):
    _fields_ = [("row_indices", C.c_void_p), ("col_indices", C.c_void_p), ("values", C.c_void_p),
                ("triplet_count", C.c_int64), ("block_dim", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [
        ("fem_dofs", C.c_int32), ("abd_dofs", C.c_int32), ("active_fem_dofs", C.c_int32), ("rank", C.c_int32),
        ("aca_relative_error", C.c_double), ("pcg_iterations", C.c_int32),
        ("schur_relative_residual", C.c_double), ("full_relative_residual", C.c_double),
        ("setup_ms", C.c_double), ("total_ms", C.c_double),
        ("active_mesh_contacts", C.c_int32), ("active_ground_contacts", C.c_int32),
        ("adaptive_abd_inverse", C.c_int32), ("neumann_doublings", C.c_int32),
        ("abd_inverse_residual", C.c_double), ("full_residual_limit", C.c_double),
        ("accepted", C.c_int32), ("fallback", C.c_int32), ("status", C.c_int32),
        ("precision", C.c_int32), ("preconditioner", C.c_int32), ("true_residual_checks", C.c_int32),
        ("restarts", C.c_int32), ("pcg_ms", C.c_double), ("descent", C.c_double),
        ("gbps", C.c_double), ("roofline_fraction", C.c_double), ("fp64_oracle_error", C.c_double),
        ("learned_used", C.c_int32), ("fallback_reason", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class SetupInfo(C.Structure):
    _fields_ = [
        ("fem_dofs", C.c_int32), ("abd_dofs", C.c_int32), ("fem_nodes", C.c_int32), ("nnzb", C.c_int32),
        ("active_fem_dofs", C.c_int32), ("rank", C.c_int32), ("lanes_per_row", C.c_int32),
        ("workgroups", C.c_int32), ("neumann_doublings", C.c_int32), ("abd_block_diagonal", C.c_int32),
        ("aca_relative_error", C.c_double), ("abd_inverse_residual", C.c_double),
        ("schedule", C.c_int32), ("pass_sync", C.c_int32), ("lambda_max", C.c_double),
        ("poly_ratio", C.c_double), ("learned_out", C.c_double * 3), ("learned_used", C.c_int32),
        ("fallback_reason", C.c_int32), ("pipelined", C.c_int32), ("pass_pairs", C.c_int32),
        ("streaming_operators", C.c_int32), ("row_tiles", C.c_int32),
    ]

    def as_dict(self) -> dict:
        out = {name: getattr(self, name) for name, _ in self._fields_}
        out["learned_out"] = list(self.learned_out)
        return out


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "hsc_abi_version": (C.c_int32, []),
    "hsc_config_default": (None, [C.POINTER(Config)]),
    "hsc_create": (C.c_int32, [C.POINTER(Config), C.c_int32, C.POINTER(C.c_void_p)]),
    "hsc_destroy": (None, [C.c_void_p]),
    "hsc_last_error": (C.c_char_p, [C.c_void_p]),
    "hsc_set_config": (C.c_int32, [C.c_void_p, C.POINTER(Config)]),
    "hsc_setup": (C.c_int32, [C.c_void_p, C.POINTER(BCOO3View), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_void_p]),
    "hsc_solve": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "hsc_do_solve": (C.c_int32, [C.c_void_p, C.POINTER(BCOO3View), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                 C.c_void_p, C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "hsc_append_jsonl": (C.c_int32, [C.POINTER(Stats), C.c_char_p, C.c_int32, C.c_int32]),
    "hsc_preconditioner_name": (C.c_char_p, [C.c_int32]),
    "hsc_precision_name": (C.c_char_p, [C.c_int32]),
    "hsc_tune_file": (C.c_int32, [C.c_char_p]),
    "hsc_tune": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "hsc_alloc_count": (C.c_int64, []),
    "hsc_solve_batched": (C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.POINTER(Stats), C.c_void_p]),
    "hsc_batch_create": (C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(C.c_void_p)]),
    "hsc_batch_launch": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "hsc_batch_stats": (C.c_int32, [C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "hsc_batch_info": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "hsc_batch_destroy": (None, [C.c_void_p]),
    "hsc_apply_D": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hsc_apply_schur": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hsc_apply_precond": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hsc_apply_pass": (C.c_int32, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "hsc_get_setup_info": (C.c_int32, [C.c_void_p, C.POINTER(SetupInfo)]),
    "hsc_fem_hessian": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "hsc_abd_fem_coupling": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                         C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hsc_assemble_bcoo": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "hsc_assemble_bcoo_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64),
                                              C.c_void_p]),
    "hsc_time_kernel": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                    C.POINTER(C.c_double)]),
    "hsc_get_factors": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "hsc_set_learned_weights": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    "hsc_get_node_graph": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "hsc_get_learned_features": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "hsc_mlp_param_count": (C.c_int32, []),
    "hsc_mlp_create": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                   C.POINTER(C.c_void_p)]),
    "hsc_mlp_destroy": (None, [C.c_void_p]),
    "hsc_mlp_set_params": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32]),
    "hsc_mlp_get_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "hsc_mlp_forward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "hsc_mlp_gradients": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                      C.POINTER(C.c_float), C.c_void_p]),
    "hsc_mlp_train_step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float,
                                       C.POINTER(C.c_float), C.c_void_p]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"HSC HIP library not built: {p} is missing "
                           "(run `python -m hsc_reimpl_amd.build`); there is no CPU fallback")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.hsc_abi_version() != HSC_ABI_VERSION:
        raise RuntimeError("HSC ABI version mismatch")
    if path is None:
        _lib = lib
    return lib
