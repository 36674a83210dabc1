"""Mixed-precision form (FP32-stored residual, FP32 MFMA) against the FP64 oracle (GPU).

Bars:
* FP32 kernels: within the worst-case FP32 dot-product bound gamma_K = K u32
  (u32 = 2^-24) of the FP64 product of the same FP32 data;
* FP32 input with FP64 arithmetic: bit-identical skeleton, ranks and draws to
  the oracle run on the FP32 values widened to FP64;
* mixed ID: its FP64-recomputed error ||A(:,rd) - A(:,J)X||_F / ||A||_F meets the
  tolerance within 10%, its rank is within one block of the oracle's, and its
  error is within 10% of the FP64 oracle's error or below the tolerance; each
  test prints both errors.
"""

import ctypes

import numpy as np
import pytest

from oracle import id_error, rbrp_colid as oracle_colid
from tests.parity import assert_skeleton
from robust_blockwise_random_pivoting_amd.matrices import (exp_spectrum, gmm_rows, kernel_block,
                                                            spectrum_matrix)

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def _cm32(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float32).T)).cuda().t()


def _dev(A):
    import torch
    return torch.as_tensor(np.ascontiguousarray(A.T)).cuda().t()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("M,N,K,nsplit", [(1, 1, 32, 1), (17, 5, 64, 3), (64, 64, 256, 4),
                                          (100, 130, 1024, 7), (29, 64, 4096, 16), (64, 300, 50048, 40),
                                          (128, 300, 4096, 8), (65, 129, 2048, 3), (300, 40, 512, 2),
                                          (128, 1000, 100000, 48)])
def test_gemm_tn32_within_fp32_bound(librbrp, hooks, M, N, K, nsplit):
    import torch
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((K, M)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    dA, dB = _cm32(A), _cm32(B)
    dC = torch.zeros((N, M), dtype=torch.float64, device="cuda").t()
    code = hooks.rbrp_test_gemm_tn32(_stream(), M, N, K, _ptr(dA), _ptr(dB), K, _ptr(dC), M, nsplit)
    assert code == 0
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64.T @ B64
    bound = (K + 2) * U32 * (np.abs(A64).T @ np.abs(B64))
    err = np.abs(dC.cpu().numpy() - ref)
    assert np.all(err <= bound), float((err / bound).max())
    # not an FP64 product in disguise nor a dropped term: typical error is FP32-sized
    assert err.max() > 0.0 or K <= 32


@pytest.mark.parametrize("M,N,K", [(128, 32, 16), (200, 70, 33), (6, 3, 1), (1000, 130, 64),
                                   (300, 17, 128), (64, 1, 100), (5000, 300, 64), (130, 33, 97),
                                   (70000, 40, 60), (50048, 300, 64), (640, 257, 33), (64, 1, 7),
                                   (1024, 129, 64), (4096, 96, 1), (640, 200, 128), (1024, 129, 97),
                                   (50048, 64, 128), (128, 33, 65)])
def test_gemm_trail32_within_fp32_bound(librbrp, hooks, M, N, K):
    """The FP32 update launcher against an FP64 product: shapes with M a multiple
    of 64 take the 128-column kernel with A shared through LDS (K <= 64 and
    K <= 128 forms), the others the 32-column kernel."""
    import torch
    rng = np.random.default_rng(M * 3 + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    dA, dC = _cm32(A), _cm32(C0)
    dBt = torch.as_tensor(np.ascontiguousarray(B)).cuda()   # row-major (k, c) at k*N + c
    ntile = (M + 63) // 64      # at most one partial row per 64 rows
    dP = torch.zeros(ntile * N, dtype=torch.float64, device="cuda")
    code = hooks.rbrp_test_gemm_upd32(_stream(), M, N, K, _ptr(dA), M, _ptr(dBt), N, _ptr(dC), M,
                                        _ptr(dP))
    assert code == 0
    A64, B64, C64 = A.astype(np.float64), B.astype(np.float64), C0.astype(np.float64)
    ref = C64 - A64 @ B64
    got = dC.cpu().numpy().astype(np.float64)
    bound = (K + 3) * U32 * (np.abs(C64) + np.abs(A64) @ np.abs(B64))
    assert np.all(np.abs(got - ref) <= bound)
    parts = dP.cpu().numpy().reshape(ntile, N)
    exact = np.sum(got * got, axis=0)
    np.testing.assert_allclose(parts.sum(axis=0), exact, rtol=64 * U32, atol=1e-300)


def test_trail32_rejects_odd_heights(librbrp, hooks):
    import torch
    z = torch.zeros(16, dtype=torch.float32, device="cuda")
    assert hooks.rbrp_test_gemm_upd32(_stream(), 5, 1, 1, _ptr(z), 6, _ptr(z), 1, _ptr(z), 6, None) == -2
    assert hooks.rbrp_test_gemm_tn32(_stream(), 1, 1, 16, _ptr(z), _ptr(z), 16, _ptr(z), 1, 1) == -4


def test_fp32_input_fp64_loop_matches_oracle():
    """rbrp_colid_s in FP64 arithmetic == the oracle on the widened FP32 values."""
    from robust_blockwise_random_pivoting_amd import rbrp_colid
    A = spectrum_matrix(600, 400, exp_spectrum(400, 15.0), 21).astype(np.float32)
    A64 = A.astype(np.float64)
    g = rbrp_colid(_cm32(A), b=32, tol=1e-6, seed=3, precision="fp64")
    r = oracle_colid(A64, 32, 1e-6, seed=3)
    perm = g.perm.cpu().numpy()
    if not assert_skeleton(perm, g.k, g.ranks, g.draws, r, "fp32 input, fp64 loop"):
        return
    eg = id_error(A64, perm, g.k, g.X.cpu().numpy())
    er = id_error(A64, r.perm, r.k, r.X)
    assert abs(eg - er) <= 0.01 * er + 1e-15


def _mixed_vs_oracle(A, b, tol, seed=7, name=""):
    from robust_blockwise_random_pivoting_amd import rbrp_colid
    g = rbrp_colid(_dev(A), b=b, tol=tol, seed=seed, precision="fp32")
    r = oracle_colid(A, b, tol, seed=seed)
    perm = g.perm.cpu().numpy()
    n = A.shape[1]
    assert np.array_equal(np.sort(perm), np.arange(n)), name
    X = g.X.cpu().numpy() if (g.k and g.k < n) else np.zeros((g.k, n - g.k))
    e32 = id_error(A, perm, g.k, X) if g.k < n else 0.0
    e64 = id_error(A, r.perm, r.k, r.X) if r.k < n else 0.0
    same = int(np.sum(perm[: min(g.k, r.k)] == r.perm[: min(g.k, r.k)]))
    print(f"[mixed {name}] k32={g.k} k64={r.k} err32={e32:.4e} err64={e64:.4e} "
          f"tracked32={g.err_final:.4e} same-prefix={same}")
    assert abs(g.k - r.k) <= b, name
    assert e32 <= 1.01 * tol, (name, e32)
    # the same error as the FP64 oracle within 10% (the storage schedule keeps
    # the FP32 rounding noise at a few percent of tol)
    assert abs(e32 - e64) <= 0.1 * e64 + 1e-13, (name, e32, e64)
    # the reported error agrees with the recomputed one
    assert g.err_final == pytest.approx(e32, rel=0.02, abs=1e-12), name
    return g, r


MIXED_CASES = [
    ("exp 300x200 b16", lambda: spectrum_matrix(300, 200, exp_spectrum(200, 10.0), 5), 16, 1e-5),
    ("exp 1000x700 b64", lambda: spectrum_matrix(1000, 700, exp_spectrum(700, 30.0), 9), 64, 1e-6),
    ("gmm 400x120 b16 full rank", lambda: gmm_rows(400, 120, 7, components=8), 16, 1e-3),
    ("kernel 200x150 b16", lambda: kernel_block(200, 150, np.random.default_rng(8)), 16, 1e-5),
    ("exp 700x500 b128", lambda: spectrum_matrix(700, 500, exp_spectrum(500, 40.0), 10), 128, 1e-6),
]


@pytest.mark.parametrize("name,gen,b,tol", MIXED_CASES)
def test_mixed_error_against_fp64_oracle(name, gen, b, tol):
    _mixed_vs_oracle(gen(), b, tol, name=name)


def test_mixed_config1():
    """1000 x 1000, s_i = exp(-i/20), b = 40, tol = 1e-6 in the mixed form."""
    A = spectrum_matrix(1000, 1000, exp_spectrum(1000, 20.0), 1)
    _mixed_vs_oracle(A, 40, 1e-6, seed=42, name="config1")


def test_mixed_gmm_robustness_small():
    """Gaussian-mixture rows with a x50 outlier component (near-full-rank run at tol 1e-6)."""
    A = gmm_rows(3000, 300, 11, components=20, outlier_frac=0.001, outlier_scale=50.0)
    g, r = _mixed_vs_oracle(A, 64, 1e-6, seed=5, name="gmm 3000x300")
    assert g.k >= 0.9 * A.shape[1]


def test_mixed_fp32_input_equals_widened_input():
    """rbrp_colid_s(A32) and rbrp_colid_d(fp64(A32)) store the same FP32 residual: identical results."""
    from robust_blockwise_random_pivoting_amd import rbrp_colid
    A = spectrum_matrix(500, 300, exp_spectrum(300, 12.0), 4).astype(np.float32)
    g1 = rbrp_colid(_cm32(A), b=24, tol=1e-5, seed=9, precision="fp32")
    g2 = rbrp_colid(_dev(A.astype(np.float64)), b=24, tol=1e-5, seed=9, precision="fp32")
    assert g1.k == g2.k
    assert np.array_equal(g1.perm.cpu().numpy(), g2.perm.cpu().numpy())
    assert np.array_equal(g1.X.cpu().numpy(), g2.X.cpu().numpy())
    assert np.array_equal(g1.error, g2.error)


def test_mixed_rejects_unknown_precision(librbrp):
    from robust_blockwise_random_pivoting_amd import Handle, _capi
    import torch
    A = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    perm = torch.zeros(8, dtype=torch.int32, device="cuda")
    with Handle() as h:
        opts = _capi.RbrpOpts()
        librbrp.rbrp_opts_default(ctypes.byref(opts))
        opts.precision = 7
        res = _capi.RbrpResult()
        res.perm = perm.data_ptr()
        assert librbrp.rbrp_colid_d(h.ptr, _ptr(A), 8, 8, 8, ctypes.byref(opts), ctypes.byref(res)) == -6


def test_mixed_large_exp_properties():
    """50000 x 4000, s_i = exp(-i/100), b = 64 at tol 1e-6: mixed and FP64 GPU runs
    reach the same error within 10% and ranks within one block (the FP64 recompute
    decides the stop where the FP32-stored sum cannot), and the mixed run's
    reported error is the recomputed one."""
    import torch
    from robust_blockwise_random_pivoting_amd import rbrp_colid
    from robust_blockwise_random_pivoting_amd.matrices import spectrum_matrix_torch
    m, n = 50000, 4000
    A = spectrum_matrix_torch(m, n, exp_spectrum(n, 100.0), seed=2024)
    g64 = rbrp_colid(A, b=64, tol=1e-6, seed=1)
    g32 = rbrp_colid(A, b=64, tol=1e-6, seed=1, precision="fp32")
    nA = torch.linalg.norm(A)

    def err(g):
        J = g.perm[: g.k].long()
        rd = g.perm[g.k:].long()
        return float(torch.linalg.norm(A[:, rd] - A[:, J] @ g.X) / nA)

    e64, e32 = err(g64), err(g32)
    print(f"[mixed 50000x4000 tol 1e-6] k64={g64.k} k32={g32.k} err64={e64:.4e} err32={e32:.4e} "
          f"reported32={g32.err_final:.4e} rel diff {abs(e32 - e64) / e64:.3f}")
    assert abs(g32.k - g64.k) < 64
    assert e64 <= 1e-6 * 1.01 and e32 <= 1e-6 * 1.01
    assert abs(e32 - e64) <= 0.1 * e64
    assert g32.err_final == pytest.approx(e32, rel=0.02)
