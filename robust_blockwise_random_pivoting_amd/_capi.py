"""ctypes binding of the C ABI declared in ``include/rbrp.h`` (``librbrp.so``) and of the
test hooks of ``include/rbrp_testing.h`` (``librbrp_testing.so``).

Only plain pointers and integers cross this boundary.  Loading fails loudly
(ImportError) when ``librbrp.so`` has not been built: there is no fallback
path.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_int, c_int32, c_int64,
                    c_uint64, c_void_p)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librbrp.so")
TESTING_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librbrp_testing.so")

RBRP_ERR_NUMERIC = 100000
RBRP_ERR_ALLOC = 100001
RBRP_ERR_SCRATCH = 100002
NOALLOC_SCRATCH_BYTES = 64 << 20
FILTERS = {"sampled": 0, "nominal": 1}
PRECISIONS = {"fp64": 0, "fp32": 1}

PHASES = ["norms", "sample", "cgs", "panel", "permute", "project", "project_reduce", "update",
          "resid", "solve"]


class RbrpOpts(Structure):
    _fields_ = [("b", c_int64), ("tol", c_double), ("max_rank", c_int64), ("seed", c_uint64),
                ("stream_offset", c_uint64), ("precision", c_int32), ("filter", c_int32)]


class RbrpResult(Structure):
    _fields_ = [("perm", c_void_p), ("X", c_void_p), ("ldx", c_int64), ("trace_cap", c_int64),
                ("ranks", c_void_p), ("error", c_void_p), ("t_ranks", c_void_p), ("k", c_int64),
                ("nblocks", c_int64), ("t_total", c_double), ("err_final", c_double),
                ("draws", c_uint64)]


class RbrpProblem(Structure):
    _fields_ = [("A", c_void_p), ("m", c_int64), ("n", c_int64), ("lda", c_int64),
                ("seed", c_uint64), ("stream_offset", c_uint64), ("perm", c_void_p),
                ("X", c_void_p), ("ldx", c_int64)]


class RbrpProblemResult(Structure):
    _fields_ = [("k", c_int64), ("nblocks", c_int64), ("err_final", c_double), ("draws", c_uint64),
                ("status", c_int32), ("pad", c_int32)]


BATCH_MODES = {"auto": 0, "persistent": 1, "loop": 2}


class RbrpPhaseStats(Structure):
    _fields_ = [("ms", c_double * len(PHASES)), ("launches", c_int64 * len(PHASES)),
                ("flops", c_double * len(PHASES)), ("bytes", c_double * len(PHASES))]


_lib = None
_hooks = None


def _hook_signatures():
    vp, i64, u64, i32p = c_void_p, c_int64, c_uint64, c_void_p
    return {
        "rbrp_test_gemm_tn": (c_int, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, c_int]),
        "rbrp_test_gemm_nn": (c_int, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, c_int, vp]),
        "rbrp_test_gemm_upd": (c_int, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp]),
        "rbrp_test_gemm_tn64": (c_int, [vp, i64, i64, i64, vp, vp, i64, vp, i64, c_int]),
        "rbrp_test_gemm_nn64": (c_int, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, c_int]),
        "rbrp_test_gram": (c_int, [vp, i64, i64, vp, i64, vp, i64, c_int]),
        "rbrp_test_trsm": (c_int, [vp, i64, i64, vp, i64, vp, i64, c_int]),
        "rbrp_test_resident": (c_int, [c_int]),
        "rbrp_test_gemm_tn32": (c_int, [vp, i64, i64, i64, vp, vp, i64, vp, i64, c_int]),
        "rbrp_test_gemm_upd32": (c_int, [vp, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp]),
        "rbrp_test_panel": (c_int, [vp, vp, i64, i64, c_int, i64, i32p, i32p, vp, i32p, i32p,
                                    POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
        "rbrp_test_panel_inv": (c_int, [vp, vp, i64, i64, c_int, i64, i32p, i32p, vp, vp, i32p,
                                        i32p, POINTER(c_int32), POINTER(c_int32),
                                        POINTER(c_int32)]),
        "rbrp_test_panel_kernel": (c_int, [vp, vp, i64, i64, c_int, i64, i32p, i32p, vp, vp,
                                           i32p, i32p, POINTER(c_int32), POINTER(c_int32),
                                           POINTER(c_int32), c_int]),
        "rbrp_test_chol2": (c_int, [vp, vp, i64, i64, vp, vp, vp, vp, POINTER(c_int32)]),
        "rbrp_test_chol2_kernel": (c_int, [vp, vp, i64, i64, vp, vp, vp, vp, POINTER(c_int32),
                                           c_int]),
        "rbrp_test_sample": (c_int, [vp, vp, i64, i64, i64, i64, u64, u64, u64, i32p,
                                     POINTER(c_int32), POINTER(c_uint64)]),
        "rbrp_test_sample_time": (c_int, [vp, vp, i64, i64, u64, c_int, POINTER(c_double)]),
        "rbrp_test_ceiling": (c_int, [vp, c_int, i64, c_int, POINTER(c_double), POINTER(c_double)]),
    }


def testing_lib():
    """Load the kernel test hooks ``librbrp_testing.so`` (include/rbrp_testing.h) once.

    Separate from ``librbrp.so`` so the product library exports only the
    include/rbrp.h surface; raises ImportError if it is missing."""
    global _hooks
    if _hooks is not None:
        return _hooks
    lib()
    if not os.path.exists(TESTING_LIB_PATH):
        raise ImportError(f"{TESTING_LIB_PATH} is not built; run build_lib")
    L = ctypes.CDLL(TESTING_LIB_PATH)
    for name, (res, args) in _hook_signatures().items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _hooks = L
    return L


def lib():
    """Load ``librbrp.so`` once; raise ImportError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m robust_blockwise_random_pivoting_amd.build_lib`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, u64, i32p = c_void_p, c_int64, c_uint64, c_void_p
    sig = {
        "rbrp_version": (c_char_p, []),
        "rbrp_opts_default": (None, [POINTER(RbrpOpts)]),
        "rbrp_philox4x64_10_host": (None, [POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
        "rbrp_create": (c_int, [POINTER(c_void_p), vp]),
        "rbrp_destroy": (c_int, [vp]),
        "rbrp_set_stream": (c_int, [vp, vp]),
        "rbrp_set_timing": (c_int, [vp, c_int]),
        "rbrp_get_phase_stats": (c_int, [vp, POINTER(RbrpPhaseStats)]),
        "rbrp_reset_phase_stats": (c_int, [vp]),
        "rbrp_release_workspace": (c_int, [vp]),
        "rbrp_uniforms_d": (c_int, [vp, u64, u64, u64, i64, vp]),
        "rbrp_colid_d": (c_int, [vp, vp, i64, i64, i64, POINTER(RbrpOpts), POINTER(RbrpResult)]),
        "rbrp_colid_s": (c_int, [vp, vp, i64, i64, i64, POINTER(RbrpOpts), POINTER(RbrpResult)]),
        "rbrp_colid_d_noalloc": (c_int, [vp, vp, i64, i64, i64, POINTER(RbrpOpts),
                                         POINTER(RbrpResult)]),
        "rbrp_colid_batched_d": (c_int, [vp, i64, POINTER(RbrpProblem), POINTER(RbrpOpts),
                                         POINTER(RbrpProblemResult), c_int]),
        "rbrp_batched_tune": (c_int, [vp, i64, POINTER(RbrpProblem), POINTER(RbrpOpts),
                                      POINTER(c_int)]),
        "rbrp_batched_launch_shape": (c_int, [vp, i64, POINTER(RbrpProblem), POINTER(RbrpOpts),
                                              POINTER(c_int), POINTER(c_int)]),
        "rbrp_gaussian_sketch_d": (c_int, [vp, u64, u64, i64, i64, c_double, vp, i64]),
        "rbrp_randcpqr_os_d": (c_int, [vp, vp, i64, i64, i64, i64, u64, u64, vp, i64, vp, vp, vp,
                                       i64]),
        "rbrp_randlupp_os_d": (c_int, [vp, vp, i64, i64, i64, i64, u64, u64, vp, i64, vp, vp, vp,
                                       i64]),
        "rbrp_randcpqr_d": (c_int, [vp, vp, i64, i64, i64, i64, u64, u64, vp, i64, vp, vp, vp, i64]),
        "rbrp_randlupp_d": (c_int, [vp, vp, i64, i64, i64, i64, u64, u64, vp, i64, vp, vp, vp, i64]),
        "rbrp_interp_os_d": (c_int, [vp, vp, i64, i64, i64, i64, vp, vp, i64]),
        "rbrp_randadaplupp_d": (c_int, [vp, vp, i64, i64, i64, c_double, i64, u64, u64, vp, vp, vp,
                                        i64, POINTER(c_int64), POINTER(c_double), POINTER(c_double)]),
        "rbrp_cpqr_id_d": (c_int, [vp, vp, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp]),
        "rbrp_matrix_ge_h": (c_int, [i64, u64, vp]),
        "rbrp_matrix_snn_h": (c_int, [i64, c_double, u64, vp]),
        "rbrp_matrix_gmm_h": (c_int, [i64, i64, u64, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class RBRPError(RuntimeError):
    """Non-zero status from the C ABI."""

    def __init__(self, fn: str, code: int):
        if code < 0:
            what = f"argument {-code} is invalid"
        elif code == RBRP_ERR_NUMERIC:
            what = "non-finite input or numerical breakdown"
        elif code == RBRP_ERR_ALLOC:
            what = "device allocation failed"
        elif code == RBRP_ERR_SCRATCH:
            what = "the problem's workspace exceeds the static scratch"
        else:
            what = f"HIP error {code}"
        super().__init__(f"{fn} returned {code}: {what}")
        self.code = code


def check(fn: str, code: int):
    if code != 0:
        raise RBRPError(fn, code)


def philox_block(ctr, key):
    """Philox4x64-10 of a 256-bit counter and 128-bit key (host, no GPU)."""
    c = (c_uint64 * 4)(*[int(x) for x in ctr])
    k = (c_uint64 * 2)(*[int(x) for x in key])
    out = (c_uint64 * 4)()
    lib().rbrp_philox4x64_10_host(c, k, out)
    return [int(x) for x in out]


__all__ = ["lib", "testing_lib", "check", "RBRPError", "RbrpOpts", "RbrpResult", "RbrpPhaseStats", "PHASES",
           "FILTERS", "PRECISIONS", "philox_block", "byref", "LIB_PATH"]
