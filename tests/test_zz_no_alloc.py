"""hsc_assemble_bcoo_noalloc: the assembly before the path without any
allocation (its workspace is the library's per-GPU device array), through
the C ABI.  Its output equals hsc_assemble_bcoo's bit for bit on two calls
in a row, the second with a smaller input; it makes no allocation (the
library's allocation counter does not move); sizes past its limit are
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from hsc_reimpl_amd import _abi, plugin, systems  # noqa: E402
from tests.test_assembly import _general_form  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _call(fn, R, C_, V, fold):
    n = int(R.numel())
    orow = torch.full((max(n, 1),), -7, dtype=torch.int32, device=DEV)
    ocol = torch.full_like(orow, -7)
    oval = torch.full((max(n, 1), 9), float("nan"), dtype=torch.float64, device=DEV)
    cnt = C.c_int64(-1)
    code = fn(C.c_void_p(R.data_ptr()), C.c_void_p(C_.data_ptr()), C.c_void_p(V.data_ptr()), n, int(fold),
              C.c_void_p(orow.data_ptr()), C.c_void_p(ocol.data_ptr()), C.c_void_p(oval.data_ptr()), C.byref(cnt),
              None)
    assert code == _abi.HSC_OK, code
    k = cnt.value
    torch.cuda.synchronize()
    return orow[:k].cpu().numpy(), ocol[:k].cpu().numpy(), oval[:k].cpu().numpy()


def test_noalloc_assembly_matches_allocating_entry_point():
    lib = _abi.load()
    big = systems.torus_system(seed=1, n_tori=3, n_bodies=3, youngs=(1e10, 1e8, 1e8))
    small = systems.cube_system(seed=0)
    cases = []
    for s, seed, both in ((big, 5, True), (small, 6, False)):
        R, Cc, V = _general_form(s, np.random.default_rng(seed), both)
        cases.append((_dev(R, torch.int32), _dev(Cc, torch.int32), _dev(V, torch.float64), not both))
    assert cases[0][0].numel() > 2 * cases[1][0].numel()
    assert cases[0][0].numel() <= 1 << 20
    for R, Cc, V, fold in cases:          # the larger input first, then the smaller one
        ref = _call(lib.hsc_assemble_bcoo, R, Cc, V, fold)
        n0 = lib.hsc_alloc_count()
        got = _call(lib.hsc_assemble_bcoo_noalloc, R, Cc, V, fold)
        assert lib.hsc_alloc_count() == n0
        assert len(got[0]) == len(ref[0]) > 0
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)


def test_noalloc_assembly_limits():
    lib = _abi.load()
    d = torch.zeros(16, dtype=torch.float64, device=DEV)
    p = C.c_void_p(d.data_ptr())
    n = C.c_int64()
    assert lib.hsc_assemble_bcoo_noalloc(p, p, p, (1 << 20) + 1, 1, p, p, p, C.byref(n), None) == \
        _abi.HSC_ERR_CAPACITY
    assert lib.hsc_assemble_bcoo_noalloc(p, p, p, -1, 1, p, p, p, C.byref(n), None) == _abi.HSC_ERR_INVALID_ARGUMENT
    assert lib.hsc_assemble_bcoo_noalloc(None, p, p, 4, 1, p, p, p, C.byref(n), None) == \
        _abi.HSC_ERR_INVALID_ARGUMENT
    assert lib.hsc_assemble_bcoo_noalloc(p, p, p, 0, 1, p, p, p, C.byref(n), None) == _abi.HSC_OK and n.value == 0
