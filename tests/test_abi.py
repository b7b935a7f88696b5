"""C ABI: the HIP library loads, exports every function include/hsc.h
declares, keeps the reference's config defaults and writes the reference's
hsc_stats.jsonl schema (host-only calls; no GPU needed)."""
import ctypes as C
import json
import re

import pytest

from hsc_reimpl_amd import _abi
from tests.conftest import ROOT


@pytest.fixture(scope="module")
def lib():
    if not _abi.LIB_PATH.exists():
        from hsc_reimpl_amd import build
        build.build()
    return _abi.load()


def declared_functions():
    text = (ROOT / "include" / "hsc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsc_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, n


def test_abi_version_and_defaults(lib):
    assert lib.hsc_abi_version() == 3
    c = _abi.Config()
    lib.hsc_config_default(C.byref(c))
    # integration.patch:34-55 and scene_default_config.cpp:146-151 of the reference
    assert (c.max_rank, c.neumann_order, c.max_iter, c.adaptive_abd_inverse) == (128, 5, 2000, 1)
    assert (c.abd_inverse_tol, c.full_residual_limit, c.omega, c.tol_rate) == (1e-5, 0.1, 0.7, 1e-3)
    assert c.residual_replace_every == 25 and c.check_acceptance == 1
    # the plugin's ABD inverse: block Neumann polynomial, doubled (linear_hsc.cu:263-297)
    assert c.abd_inverse == _abi.ABD_INVERSE["neumann"] and c.precision == _abi.PRECISION["fp64"]
    assert c.preconditioner == _abi.PRECONDITIONER["neumann"]
    # extensions: autotuned schedule, Chebyshev interval ratio, neighbour-tile pass sync
    assert (c.schedule, c.poly_interval_ratio, c.pass_sync) == (0, 100.0, 2)
    assert (c.pipelined, c.operator_mapping) == (0, 0)


def test_invalid_config_rejected(lib):
    c = _abi.Config()
    lib.hsc_config_default(C.byref(c))
    c.omega = 1.5
    h = C.c_void_p()
    assert lib.hsc_create(C.byref(c), 0, C.byref(h)) == _abi.HSC_ERR_INVALID_ARGUMENT
    lib.hsc_config_default(C.byref(c))
    c.neumann_order = 0
    assert lib.hsc_create(C.byref(c), 0, C.byref(h)) == _abi.HSC_ERR_INVALID_ARGUMENT
    for field, bad in (("schedule", 4), ("pass_sync", 3), ("pass_sync", -1), ("lanes_per_row", 6),
                       ("operator_mapping", 3), ("operator_mapping", -1)):
        lib.hsc_config_default(C.byref(c))
        setattr(c, field, bad)
        assert lib.hsc_create(C.byref(c), 0, C.byref(h)) == _abi.HSC_ERR_INVALID_ARGUMENT, field
    lib.hsc_config_default(C.byref(c))
    c.preconditioner = _abi.PRECONDITIONER["chebyshev"]
    c.poly_interval_ratio = 1.0
    assert lib.hsc_create(C.byref(c), 0, C.byref(h)) == _abi.HSC_ERR_INVALID_ARGUMENT
    assert lib.hsc_solve(None, None, None, None, None) == _abi.HSC_ERR_INVALID_ARGUMENT
    ms = C.c_double()
    assert lib.hsc_time_kernel(None, 0, 8, 1, None, C.byref(ms)) == _abi.HSC_ERR_INVALID_ARGUMENT
    assert lib.hsc_apply_pass(None, 0, 0.5, None, None, None, None) == _abi.HSC_ERR_INVALID_ARGUMENT


def test_stateless_entry_points_validate_before_any_device_call(lib):
    """hsc_fem_hessian, hsc_assemble_bcoo and hsc_abd_fem_coupling reject
    bad sizes and null pointers, and sizes past their limits, before they
    touch the device (the dummy pointer is never dereferenced)."""
    buf = (C.c_double * 16)()
    d = C.cast(buf, C.c_void_p)
    INV, CAP = _abi.HSC_ERR_INVALID_ARGUMENT, _abi.HSC_ERR_CAPACITY
    # FEM generator: negative count, nulls with a positive count, past 2^27 tets
    assert lib.hsc_fem_hessian(d, d, d, d, d, d, 1.0, -1, 1, d, d, d, None) == INV
    assert lib.hsc_fem_hessian(None, d, d, d, d, d, 1.0, 5, 1, d, d, d, None) == INV
    assert lib.hsc_fem_hessian(d, d, d, d, d, d, 1.0, (1 << 27) + 1, 1, d, d, d, None) == CAP
    # assembly: no count pointer, negative count, nulls, past 2^31 - 1 triplets
    n = C.c_int64()
    assert lib.hsc_assemble_bcoo(d, d, d, 4, 1, d, d, d, None, None) == INV
    assert lib.hsc_assemble_bcoo(d, d, d, -1, 1, d, d, d, C.byref(n), None) == INV
    assert lib.hsc_assemble_bcoo(None, d, d, 4, 1, d, d, d, C.byref(n), None) == INV
    assert lib.hsc_assemble_bcoo(d, d, d, 1 << 31, 1, d, d, d, C.byref(n), None) == CAP
    # coupling: negative sizes / offsets, nulls, past 2^28 contacts, block indices past int32
    args = dict(av=d, fv=d, h=d, count=4, v2b=d, xb=d, nv=8, bf=d, nb=2, vf=d, nf=10, ao=0, fo=8)

    def cpl(**kw):
        a = dict(args, **kw)
        return lib.hsc_abd_fem_coupling(a["av"], a["fv"], a["h"], a["count"], a["v2b"], a["xb"], a["nv"], a["bf"],
                                        a["nb"], a["vf"], a["nf"], a["ao"], a["fo"], d, d, d, None)
    for kw in ({"count": -1}, {"nv": -1}, {"nb": -1}, {"nf": -1}, {"ao": -1}, {"fo": -1}, {"av": None},
               {"xb": None}, {"vf": None}):
        assert cpl(**kw) == INV, kw
    assert cpl(count=(1 << 28) + 1) == CAP
    assert cpl(nb=1 << 30) == CAP
    assert cpl(count=0) == _abi.HSC_OK          # nothing to do, no device call


def test_jsonl_row_matches_reference_schema(lib, tmp_path):
    s = _abi.Stats()
    s.fem_dofs, s.abd_dofs, s.active_fem_dofs, s.rank = 74508, 144, 819, 84
    s.pcg_iterations, s.schur_relative_residual, s.full_relative_residual = 499, 9.8e-5, 1.4e-7
    s.accepted, s.adaptive_abd_inverse, s.full_residual_limit = 1, 1, 0.1
    s.active_mesh_contacts, s.active_ground_contacts = 1234, 56
    s.preconditioner, s.precision = _abi.PRECONDITIONER["chebyshev"], _abi.PRECISION["mixed"]
    s.pcg_ms, s.gbps, s.roofline_fraction, s.fp64_oracle_error = 3.5, 2500.0, 0.3125, -1.0
    p = tmp_path / "hsc_stats.jsonl"
    assert lib.hsc_append_jsonl(C.byref(s), str(p).encode(), 1, 0) == 0
    s.fp64_oracle_error, s.preconditioner, s.learned_used = 2.5e-9, _abi.PRECONDITIONER["learned"], 1
    assert lib.hsc_append_jsonl(C.byref(s), str(p).encode(), 1, 1) == 0
    rows = [json.loads(line) for line in p.read_text().splitlines()]
    # field list of linear_hsc.cu:421-431 and the consumer storage.py:48-61 of the reference
    ref_keys = ["frame", "newton", "fem_dofs", "abd_dofs", "active_fem_dofs", "rank", "aca_relative_error",
                "pcg_iterations", "schur_relative_residual", "full_relative_residual", "setup_ms", "total_ms",
                "active_mesh_contacts", "active_ground_contacts", "adaptive_abd_inverse", "neumann_doublings",
                "abd_inverse_residual", "full_residual_limit", "accepted", "solver", "fallback"]
    assert list(rows[0].keys())[:len(ref_keys)] == ref_keys
    assert rows[1]["newton"] == 1 and rows[0]["solver"] == "linear_hsc" and rows[0]["fallback"] is False
    assert rows[0]["accepted"] is True and rows[0]["adaptive_abd_inverse"] is True
    # the consumer's acceptance rule
    assert not (rows[0]["solver"] != "linear_hsc" or rows[0]["fallback"] or not rows[0]["accepted"])
    # contact counters supplied by the caller, names of what ran, rate and oracle-error fields
    assert (rows[0]["active_mesh_contacts"], rows[0]["active_ground_contacts"]) == (1234, 56)
    assert rows[0]["preconditioner"] == "chebyshev" and rows[0]["precision"] == "mixed"
    assert rows[1]["preconditioner"] == "learned" and rows[1]["learned_used"] is True
    assert rows[0]["gbps"] == 2500.0 and rows[0]["roofline_fraction"] == 0.3125
    assert rows[0]["fp64_oracle_error"] is None and rows[1]["fp64_oracle_error"] == 2.5e-9


def test_names_of_modes(lib):
    names = {v: lib.hsc_preconditioner_name(v).decode() for v in _abi.PRECONDITIONER.values()}
    assert names == {v: k for k, v in _abi.PRECONDITIONER.items()}
    assert lib.hsc_preconditioner_name(7) == b"unknown"
    assert [lib.hsc_precision_name(v).decode() for v in (0, 1)] == ["fp64", "mixed"]


def test_alloc_count_and_tune_file_without_a_device(lib, tmp_path):
    # host-only calls: the counter reads, and a table file is loaded (lines
    # of another format are skipped) or an absent one is tolerated
    assert lib.hsc_alloc_count() >= 0
    f = tmp_path / "tune.txt"
    f.write_text("garbage line\nhsc-tune-v4 256 14 11 3 1 0 1 16 1 4 2 1\n")
    assert lib.hsc_tune_file(str(f).encode()) == 0
    assert lib.hsc_tune_file(str(tmp_path / "absent.txt").encode()) == 0
    assert lib.hsc_tune_file(None) == 0
    assert lib.hsc_tune(None, None) == _abi.HSC_ERR_INVALID_ARGUMENT
