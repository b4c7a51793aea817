"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, and
rejects invalid arguments on the host before touching a device."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT
from cytnx_amd import _lib


def _declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(cxa_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_loads_and_abi_version():
    L = _lib.lib()
    assert L.cxa_abi_version() == _lib.ABI_VERSION


def test_every_declared_symbol_is_exported():
    declared = _declared_functions()
    assert "cxa_permute" in declared
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (cxa_[a-z0-9_]+)", out))
    assert set(declared) <= exported
    # nothing exported under the cxa_ prefix that the header does not declare
    assert exported <= set(declared), sorted(exported - set(declared))


def test_permute_rejects_bad_mapper_on_host():
    L = _lib.lib()
    shape = _lib.u64_array([2, 3])
    rc = L.cxa_permute(3, 2, shape, _lib.u64_array([0, 0]), None, None, None)
    assert rc == _lib.ERR_INVALID
    assert b"duplicated" in L.cxa_last_error()
    rc = L.cxa_permute(3, 2, shape, _lib.u64_array([0, 2]), None, None, None)
    assert rc == _lib.ERR_INVALID
    rc = L.cxa_permute(99, 2, shape, _lib.u64_array([1, 0]), None, None, None)
    assert rc == _lib.ERR_UNSUPPORTED
    # zero elements: nothing to move, succeeds without touching memory
    rc = L.cxa_permute(3, 2, _lib.u64_array([0, 3]), _lib.u64_array([1, 0]), None, None, None)
    assert rc == _lib.OK


def test_block_permute_plan_validates_on_host():
    L = _lib.lib()
    h = ctypes.c_void_p(0)
    rc = L.cxa_block_permute_plan(3, 2, 0, None, None, None, _lib.u64_array([0, 0]),
                                  ctypes.byref(h))
    assert rc == _lib.ERR_INVALID and b"permutation" in L.cxa_last_error()
    rc = L.cxa_block_permute_plan(99, 2, 0, None, None, None, _lib.u64_array([1, 0]),
                                  ctypes.byref(h))
    assert rc == _lib.ERR_UNSUPPORTED
    rc = L.cxa_block_permute_plan(3, 2, 1, None, None, None, _lib.u64_array([1, 0]),
                                  ctypes.byref(h))
    assert rc == _lib.ERR_INVALID
    # no blocks: an empty plan, built and executed without touching the device
    rc = L.cxa_block_permute_plan(3, 2, 0, None, None, None, _lib.u64_array([1, 0]),
                                  ctypes.byref(h))
    assert rc == _lib.OK and h.value
    assert L.cxa_block_permute_exec(h, None, None, None) == _lib.OK
    assert L.cxa_block_permute_destroy(h) == _lib.OK


def test_strided_batch_plan_validates_on_host():
    L = _lib.lib()
    h = ctypes.c_void_p(0)
    ok = (3, 4, 5, 6, 7, 0, 7, 1, 0, 5, 1, 30, 5)  # dtype, nbatch, M, N, K, a, b, c strides
    bad = [(99,) + ok[1:], ok[:2] + (-1,) + ok[3:], ok[:5] + (-1,) + ok[6:],
           ok[:12] + (4,)]  # unknown dtype, negative M, negative stride, ldc < N
    for args in bad:
        rc = L.cxa_strided_batch_plan(*args, ctypes.byref(h))
        assert rc in (_lib.ERR_INVALID, _lib.ERR_UNSUPPORTED) and not h.value


def test_check_maps_status_to_exceptions():
    with pytest.raises(_lib.CxaInvalid):
        L = _lib.lib()
        _lib.check(L.cxa_permute(3, 1, _lib.u64_array([4]), _lib.u64_array([3]), None, None,
                                 None))


def test_vector_reductions_validate_on_host():
    L = _lib.lib()
    ws = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation first
    assert L.cxa_vec_workspace_bytes() >= 2 * 512 * 8
    rc = L.cxa_dot(5, 4, ctypes.c_void_p(8), ctypes.c_void_p(8), 0, ctypes.c_void_p(8), ws, None)
    assert rc == _lib.ERR_UNSUPPORTED
    rc = L.cxa_nrm2(3, -1, ctypes.c_void_p(8), ctypes.c_void_p(8), ws, None)
    assert rc == _lib.ERR_INVALID
    rc = L.cxa_dot(3, 4, None, None, 0, ctypes.c_void_p(8), ws, None)
    assert rc == _lib.ERR_INVALID


def test_tensordot_noalloc_rejects_oversize_on_host():
    """The no-allocation tensordot refuses operands whose permuted copies exceed its static
    scratch, before any device work (so no pointers are needed)."""
    L = _lib.lib()
    # L (128, 3, 16384) contracted over axis 1: its free axes (0, 2) are not adjacent, so L is
    # copied: 6,291,456 doubles = 48 MiB > 32 MiB
    rc = L.cxa_tensordot_noalloc(3, None, None, 3, _lib.u64_array([128, 3, 16384]), None, 2,
                                 _lib.u64_array([3, 2]), 1, _lib.u64_array([1]),
                                 _lib.u64_array([0]), None)
    assert rc == _lib.ERR_INVALID
    assert b"static scratch" in L.cxa_last_error()
    rc = L.cxa_tensordot_noalloc(5, None, None, 1, _lib.u64_array([2]), None, 1,
                                 _lib.u64_array([2]), 1, _lib.u64_array([0]),
                                 _lib.u64_array([0]), None)
    assert rc == _lib.ERR_UNSUPPORTED
