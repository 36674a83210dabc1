"""The reference driver's test matrices (gpu/driver.cpp:23-111) and the rbrp_driver program.

The generators run on the host through the C ABI (CPU tests); the driver
program runs RBRP, the CPQR ID and the two sketch IDs on the device (GPU test)
and prints the reference's ranks / error / time lists (gpu/driver.cpp:182-304).
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest


def _gen(librbrp, kind, *args):
    if kind == "ge":
        n, seed = args
        A = np.zeros((n, n), order="F")
        code = librbrp.rbrp_matrix_ge_h(n, seed, ctypes.c_void_p(A.ctypes.data))
    elif kind == "snn":
        n, s, seed = args
        A = np.zeros((n, n), order="F")
        code = librbrp.rbrp_matrix_snn_h(n, s, seed, ctypes.c_void_p(A.ctypes.data))
    else:
        n, d, seed = args
        A = np.zeros((d, n), order="F")
        code = librbrp.rbrp_matrix_gmm_h(n, d, seed, ctypes.c_void_p(A.ctypes.data))
    return code, A


def test_ge_has_the_prescribed_singular_values(librbrp):
    n = 260
    code, A = _gen(librbrp, "ge", n, 5)
    assert code == 0
    s = np.array([1.0 if i < 100 else max(0.8 ** (i - 100), 1e-5) for i in range(n)])
    sv = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(sv, np.sort(s)[::-1], rtol=0, atol=1e-13)
    code2, A2 = _gen(librbrp, "ge", n, 5)
    assert np.array_equal(A, A2)
    assert not np.array_equal(A, _gen(librbrp, "ge", n, 6)[1])


def test_snn_is_a_nonnegative_sparse_product(librbrp):
    n, s = 400, 0.05
    code, A = _gen(librbrp, "snn", n, s, 9)
    assert code == 0
    assert A.min() >= 0.0 and A.max() > 0.0
    # each of the n ceil(s n) entries of V adds one column of U D (about ceil(s n)
    # entries) to A, so A holds at most n ceil(s n)^2 non-zeros
    nnz = int(np.ceil(s * n))
    assert 0 < np.count_nonzero(A) <= n * nnz * nnz
    # the columns of U D carry the weights d_c: the first 100 (d = 0.1) dominate,
    # so A is numerically low-rank next to its size
    sv = np.linalg.svd(A, compute_uv=False)
    assert sv[150] < 0.05 * sv[0]
    assert np.array_equal(A, _gen(librbrp, "snn", n, s, 9)[1])


def test_gmm_cluster_shifts_and_noise(librbrp):
    n, d = 1000, 120
    code, A = _gen(librbrp, "gmm", n, d, 4)
    assert code == 0
    per = n // 100
    shift = np.zeros_like(A)
    for i in range(n):
        shift[i // per, i] = 10.0 * (i // per + 1)
    Z = A - shift
    assert abs(Z.mean()) < 5.0 / np.sqrt(Z.size)
    assert abs(Z.std() - 1.0) < 0.01
    # each cluster's shifted entries sit at 10 (j + 1) on average
    for j in (0, 37, 99):
        cols = slice(j * per, (j + 1) * per)
        assert abs(A[j, cols].mean() - 10.0 * (j + 1)) < 5.0 / np.sqrt(per)


def test_generator_argument_errors(librbrp):
    assert _gen(librbrp, "gmm", 99, 200, 0)[0] == -1
    assert _gen(librbrp, "gmm", 1000, 99, 0)[0] == -2
    assert _gen(librbrp, "gmm", 150, 120, 0)[0] == -2     # cluster rows up to 149 needed
    assert librbrp.rbrp_matrix_snn_h(10, 0.0, 0, None) == -2
    assert librbrp.rbrp_matrix_ge_h(0, 0, None) == -1


def _sections(text):
    """{section: {list name: values}} from the driver's output."""
    out, sec, lines = {}, "RBRP", text.splitlines()
    for i, ln in enumerate(lines):
        ln = ln.strip()
        if ln.startswith("[") and ln.endswith("]"):
            sec = ln[1:-1]
        elif ln in ("ranks:", "error:", "time:"):
            out.setdefault(sec, {})[ln[:-1]] = [float(x) for x in lines[i + 1].split()]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["-n", "300", "-bs", "16", "-nb", "4", "-matrix", "ge", "-seed", "3"],
                                  ["-n", "400", "-bs", "24", "-nb", "3", "-matrix", "gmm", "-d", "250"]])
def test_driver_prints_the_reference_lists(librbrp, args):
    from robust_blockwise_random_pivoting_amd import build_lib
    assert os.path.exists(build_lib.DRIVER), "build the library first"
    print("rbrp_driver", " ".join(args), flush=True)
    r = subprocess.run([build_lib.DRIVER] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Create test matrix:" in r.stdout and "RBRP total time:" in r.stdout
    assert "CPQR total time:" in r.stdout
    sec = _sections(r.stdout)
    nb = int(args[args.index("-nb") + 1])
    bs = int(args[args.index("-bs") + 1])
    ranks = sec["RBRP"]["ranks"]
    assert len(ranks) == nb and all(0 < a < b for a, b in zip(ranks, ranks[1:]))
    assert ranks[0] <= bs
    e_rbrp = sec["RBRP"]["error"]
    assert all(a >= b for a, b in zip(e_rbrp, e_rbrp[1:])) and e_rbrp[0] < 1.0
    for name in ("CPQR", "RandCPQR-OS", "RandLUPP-OS"):
        e, t = sec[name]["error"], sec[name]["time"]
        assert len(e) == nb and len(t) == nb and min(t) > 0
        # every method's squared error at a rank stays within two orders of
        # magnitude of RBRP's at that rank
        for a, b in zip(e, e_rbrp):
            assert 0 <= a <= 100 * b + 1e-14, (name, e, e_rbrp)
