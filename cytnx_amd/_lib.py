"""Loader for libcytnx_amd.so, the C ABI declared in include/cytnx_amd.h.

The library is required: every device operation in this package goes through it, and there is
no CPU or PyTorch fallback.  A missing or stale build raises immediately.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcytnx_amd.so")
ABI_VERSION = 6

OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_NOMEM, ERR_NUMERIC = 0, 1, 2, 3, 4, 5

_lib = None


class CxaError(RuntimeError):
    """A C-ABI call returned a non-zero status; ``code`` is that status."""

    def __init__(self, code: int, msg: str):
        super().__init__("[cytnx_amd status %d] %s" % (code, msg))
        self.code = code


class CxaInvalid(CxaError, ValueError):
    pass


u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p


ci = ctypes.c_int
i64 = ctypes.c_int64
dblp = ctypes.POINTER(ctypes.c_double)
intp = ctypes.POINTER(ctypes.c_int)


def _declare(L):
    sig = {
        "cxa_last_error": (ctypes.c_char_p, []),
        "cxa_abi_version": (ci, []),
        "cxa_permute": (ci, [ci, ci, u64p, u64p, vp, vp, vp]),
        "cxa_block_permute_plan": (ci, [ci, ci, i64, u64p, i64p, i64p, u64p, ctypes.c_void_p]),
        "cxa_strided_batch_plan": (ci, [ci] + [i64] * 12 + [ctypes.c_void_p]),
        "cxa_block_permute_exec": (ci, [vp, vp, vp, vp]),
        "cxa_block_permute_destroy": (ci, [vp]),
        "cxa_matmul": (ci, [ci, vp, vp, vp, i64, i64, i64, vp]),
        "cxa_gemm_batch": (ci, [ci, ctypes.c_char_p, ctypes.c_char_p, i64p, i64p, i64p, dblp,
                                ctypes.POINTER(vp), i64p, ctypes.POINTER(vp), i64p, dblp,
                                ctypes.POINTER(vp), i64p, i64, i64p, vp]),
        "cxa_tensordot_workspace": (ci, [ci, ci, u64p, ci, u64p, ci, u64p, u64p, i64p]),
        "cxa_tensordot": (ci, [ci, vp, vp, ci, u64p, vp, ci, u64p, ci, u64p, u64p, vp, vp]),
        "cxa_tensordot_noalloc": (ci, [ci, vp, vp, ci, u64p, vp, ci, u64p, ci, u64p, u64p, vp]),
        "cxa_u1_enumerate_blocks": (ci, [ci, intp, i64p, i64p, i64, u64p, i64p]),
        "cxa_block_contract_plan": (ci, [ci, ci, ci, u64p, u64p, i64p, ci, ci, u64p, u64p, i64p,
                                         ci, u64p, u64p, ci, u64p, u64p, i64p,
                                         ctypes.POINTER(vp)]),
        "cxa_block_contract_info": (ci, [vp, dblp, i64p, i64p]),
        "cxa_block_contract_set_tile": (ci, [vp, ci]),
        "cxa_block_contract_set_split": (ci, [vp, ci]),
        "cxa_block_contract_set_config": (ci, [vp, ci, ci, ci]),
        "cxa_block_contract_get_config": (ci, [vp, ci, intp, intp, intp]),
        "cxa_block_contract_exec": (ci, [vp, vp, vp, vp, vp, ci, vp]),
        "cxa_block_contract_exec_ex": (ci, [vp, vp, vp, vp, vp, ci, ci, vp]),
        "cxa_block_contract_destroy": (ci, [vp]),
        "cxa_mix_contract_plan": (ci, [ci, ci, i64p, i64, i64p, i64, i64p, dblp,
                                       ctypes.POINTER(vp)]),
        "cxa_heff2_mix_plan": (ci, [ci, i64p, i64p, i64p, intp, ci, i64p, i64p, i64p, intp, ci,
                                    i64p, i64p, i64p, intp, i64p, i64p, dblp, i64p, vp, i64, i64,
                                    i64, i64p, i64p, i64p, dblp, vp, i64p]),
        "cxa_vec_workspace_bytes": (ci, []),
        "cxa_dot": (ci, [ci, i64, vp, vp, ci, vp, vp, vp]),
        "cxa_nrm2": (ci, [ci, i64, vp, vp, vp, vp]),
        "cxa_tridiag_ground": (ci, [ci, dblp, dblp, dblp, dblp]),
        "cxa_chain_create": (ci, [ci, ctypes.POINTER(vp), ctypes.POINTER(vp), intp, i64p, i64p,
                                  ctypes.POINTER(vp)]),
        "cxa_chain_prepare": (ci, [vp, ci, vp]),
        "cxa_chain_matvec": (ci, [vp, vp, vp, vp]),
        "cxa_chain_step": (ci, [vp, ci, vp, vp, vp]),
        "cxa_chain_tune": (ci, [vp, vp, vp, ci, intp, intp, dblp, vp]),
        "cxa_chain_flops": (ci, [vp, dblp]),
        "cxa_chain_layout": (ci, [vp, ci, intp]),
        "cxa_chain_destroy": (ci, [vp]),
        "cxa_chain_graph_create": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp)]),
        "cxa_graph_launch": (ci, [vp, vp]),
        "cxa_graph_destroy": (ci, [vp]),
        "cxa_lanczos_create": (ci, [i64, ci, ctypes.POINTER(vp)]),
        "cxa_lanczos_run": (ci, [vp, vp, vp, vp, ctypes.c_double, ci, ci, ci, dblp, vp, intp,
                                 dblp, dblp, vp]),
        "cxa_lanczos_destroy": (ci, [vp]),
        "cxa_lanczos_reshape": (ci, [vp, i64, ci]),
        "cxa_lanczos_exp": (ci, [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ci, ci, vp, intp, vp]),
        "cxa_lanczos_eigsh": (ci, [vp, vp, vp, vp, ci, ci, ctypes.c_double, ci, ci, ci, dblp,
                                   vp, intp, intp, vp]),
        "cxa_tridiag_eigh": (ci, [ci, dblp, dblp, dblp, dblp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return sig


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "cytnx_amd: native library %s is missing; build it with "
                "`python -m cytnx_amd.build` (there is no fallback path)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        v = L.cxa_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError("cytnx_amd: library ABI %d, package expects %d; rebuild"
                               % (v, ABI_VERSION))
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().cxa_last_error().decode(errors="replace")
        if rc == ERR_INVALID:
            raise CxaInvalid(rc, msg)
        raise CxaError(rc, msg)


def u64_array(seq):
    seq = [int(x) for x in seq]
    return (ctypes.c_uint64 * max(1, len(seq)))(*seq)


def i64_array(seq):
    seq = [int(x) for x in seq]
    return (ctypes.c_int64 * max(1, len(seq)))(*seq)


def int_array(seq):
    seq = [int(x) for x in seq]
    return (ctypes.c_int * max(1, len(seq)))(*seq)
