"""C-ABI checks that need no GPU: the library loads, exports every declared
symbol, its host Philox matches the known-answer vectors, and argument errors
come back as status codes."""

import ctypes
import os
import re

import numpy as np

from tests.test_oracle import KAT, numpy_philox_block

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    names = set()
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    for m in re.finditer(r"\b(rbrp_\w+)\s*\(", txt):
        names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported(librbrp, hooks):
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["rbrp.h", "rbrp_testing.h"]
    names = declared_functions("rbrp.h")
    assert "rbrp_colid_d" in names and len(names) >= 10
    for name in names:
        assert hasattr(librbrp, name), name
    test_names = declared_functions("rbrp_testing.h")
    assert test_names and all(n.startswith("rbrp_test_") for n in test_names)
    for name in test_names:
        assert hasattr(hooks, name), name
        # the product library exports the rbrp.h surface only
        assert not hasattr(librbrp, name), name


def test_host_philox_kat(librbrp):
    from robust_blockwise_random_pivoting_amd._capi import philox_block
    for ctr, key, want in KAT:
        assert tuple(philox_block(ctr, key)) == want


def test_host_philox_matches_numpy_random_counters(librbrp):
    from robust_blockwise_random_pivoting_amd._capi import philox_block
    rng = np.random.default_rng(5)
    for _ in range(50):
        ctr = [int(x) for x in rng.integers(0, 2 ** 63, size=4, dtype=np.int64)]
        ctr[0] = (ctr[0] << 1) | 1
        key = [int(x) for x in rng.integers(0, 2 ** 63, size=2, dtype=np.int64)]
        assert tuple(philox_block(ctr, key)) == numpy_philox_block(ctr, key)


def test_argument_errors_are_status_codes(librbrp, hooks):
    from robust_blockwise_random_pivoting_amd import _capi
    opts = _capi.RbrpOpts()
    librbrp.rbrp_opts_default(ctypes.byref(opts))
    assert opts.b == 64 and opts.tol == 0.0
    res = _capi.RbrpResult()
    # null handle -> -1; nothing touches a device
    assert librbrp.rbrp_colid_d(None, None, 10, 10, 10, ctypes.byref(opts), ctypes.byref(res)) == -1
    assert librbrp.rbrp_destroy(None) == -1
    assert librbrp.rbrp_create(None, None) == -1
    assert librbrp.rbrp_version().decode().startswith("rbrp")
    # kernel-level hooks validate before launching
    assert hooks.rbrp_test_gemm_tn(None, 4, 4, 10, None, 16, None, 16, None, 4, 1) == -4
    assert hooks.rbrp_test_panel(None, None, 0, 8, 0, 0, None, None, None, None, None,
                                   None, None, None) == -3
