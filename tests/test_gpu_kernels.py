"""Kernel-level parity of the HIP library against host computations (GPU)."""

import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import robust_filter, sample_without_replacement, symmetric_difference, uniform_stream

pytestmark = pytest.mark.gpu


def _t(x, dtype=None):
    import torch
    return torch.as_tensor(np.asfortranarray(x) if np.ndim(x) == 2 else x, dtype=dtype).cuda()


def _cm(a):
    """Column-major device copy of a host 2-D array."""
    import torch
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda().t()


def _host(t):
    return t.cpu().numpy()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_uniforms_match_numpy_philox(librbrp):
    from robust_blockwise_random_pivoting_amd import uniforms
    for seed, off, d0, cnt in [(0, 0, 0, 1000), (42, 0, 3, 513), (2 ** 63 + 5, 2 ** 64 - 2, 1, 64)]:
        ref = uniform_stream(seed, off).random(d0 + cnt)[d0:]
        got = _host(uniforms(seed, off, d0, cnt))
        assert np.array_equal(got, ref), (seed, off, d0)


@pytest.mark.parametrize("M,N,K,nsplit", [(1, 1, 16, 1), (17, 5, 64, 3), (64, 64, 256, 4),
                                          (100, 130, 1024, 7), (29, 64, 4096, 16)])
def test_gemm_tn_matches_numpy(librbrp, hooks, M, N, K, nsplit):
    import torch
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    dA, dB = _cm(A), _cm(B)
    dC = _cm(np.zeros((M, N)))
    code = hooks.rbrp_test_gemm_tn(_stream(), M, N, K, _ptr(dA), K, _ptr(dB), K, _ptr(dC), M, nsplit)
    assert code == 0
    ref = A.T @ B
    assert np.abs(_host(dC) - ref).max() <= 1e-12 * np.abs(ref).max() * np.sqrt(K)
    torch.cuda.synchronize()


@pytest.mark.parametrize("M,N,K,nsplit", [(1, 1, 16, 1), (17, 5, 64, 3), (64, 64, 256, 4),
                                          (100, 130, 1024, 7), (29, 64, 4096, 16),
                                          (64, 300, 50048, 40), (64, 129, 512, 1)])
def test_gemm_tn64_matches_numpy(librbrp, hooks, M, N, K, nsplit):
    """The 64 x 128-tile projection kernel against NumPy at FP64 rounding."""
    import torch
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    dA, dB = _cm(A), _cm(B)
    dC = _cm(np.zeros((M, N)))
    code = hooks.rbrp_test_gemm_tn64(_stream(), M, N, K, _ptr(dA), _ptr(dB), K, _ptr(dC), M, nsplit)
    assert code == 0
    ref = A.T @ B
    assert np.abs(_host(dC) - ref).max() <= 1e-12 * np.abs(ref).max() * np.sqrt(K)
    torch.cuda.synchronize()


@pytest.mark.parametrize("M,N,K,mode", [(64, 64, 16, 0), (132, 70, 33, 1), (4, 3, 1, 2),
                                        (50048, 64, 1000, 1), (256, 17, 2047, 2), (128, 64, 48, 1),
                                        # row ranges of 1..4 blocks, rows not a multiple of 16,
                                        # two column tiles, more ranges than one per workgroup
                                        (1004, 64, 100, 1), (16, 3, 5, 0), (70000, 128, 300, 1),
                                        (20, 64, 64, 2), (300000, 64, 40, 1)])
def test_gemm_nn64_matches_numpy(librbrp, hooks, M, N, K, mode):
    """The long-reduction NN kernel against NumPy at FP64 rounding (set / subtract / add).
    Its rows are split into ranges of 16-row blocks balanced over the resident
    workgroups, so the cases cover ranges of one to four blocks, partial blocks,
    two column tiles and more ranges than resident workgroups."""
    import torch
    rng = np.random.default_rng(M + 7 * N + 11 * K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    dA, dB, dC = _cm(A), _cm(B), _cm(C0)
    code = hooks.rbrp_test_gemm_nn64(_stream(), M, N, K, _ptr(dA), M, _ptr(dB), K, _ptr(dC), M, mode)
    assert code == 0
    prod = A @ B
    ref = prod if mode == 0 else (C0 - prod if mode == 1 else C0 + prod)
    scale = np.abs(A).max() * np.abs(B).max() * K + np.abs(C0).max()
    assert np.abs(_host(dC) - ref).max() <= 1e-14 * scale
    torch.cuda.synchronize()


@pytest.mark.parametrize("N,K,nsplit", [(64, 50048, 241), (1, 16, 1), (17, 160, 3), (40, 3008, 16),
                                        (64, 64, 1), (33, 1024, 7), (48, 50048, 256),
                                        (128, 100000, 256), (65, 4000, 9), (71, 1000000, 256),
                                        (100, 336, 2)])
def test_gram_matches_numpy(librbrp, hooks, N, K, nsplit):
    """The panel's symmetric Gram kernel (64- and 128-column tiles) against NumPy
    at FP64 rounding: partial column blocks, one to many splits; the result is
    exactly symmetric."""
    import torch
    rng = np.random.default_rng(N * 7 + K)
    A = rng.standard_normal((K, N)) * np.exp(-np.arange(N) / 10.0)
    dA = _cm(A)
    dC = _cm(np.full((N, N), np.nan))
    code = hooks.rbrp_test_gram(_stream(), N, K, _ptr(dA), K, _ptr(dC), N, nsplit)
    assert code == 0
    got = _host(dC)
    ref = A.T @ A
    assert np.array_equal(got, got.T)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max() * np.sqrt(K / 16.0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("k,nc,use_dinv", [(1, 5, 1), (17, 16, 1), (64, 17, 1), (65, 100, 1),
                                           (300, 33, 1), (2047, 1953, 1), (257, 40, 0), (70, 3, 0)])
def test_trsm_matches_scipy(librbrp, hooks, k, nc, use_dinv):
    """The X solve's blocked back substitution (one-launch column tiles over the
    diagonal-block inverses, and the substitution path) against SciPy on a
    well-conditioned upper-triangular R: partial top blocks, column counts off
    the 16-column tile, ld larger than k."""
    import torch
    rng = np.random.default_rng(k * 3 + nc)
    R = np.triu(rng.standard_normal((k, k))) * 0.1
    R[np.diag_indices(k)] = 1.0 + rng.random(k)
    ldr, ldx = k + 3, k + 5
    Rp = np.zeros((ldr, k)); Rp[:k] = R
    X0 = rng.standard_normal((k, nc))
    Xp = np.zeros((ldx, nc)); Xp[:k] = X0
    dR, dX = _cm(Rp), _cm(Xp)
    code = hooks.rbrp_test_trsm(_stream(), k, nc, _ptr(dR), ldr, _ptr(dX), ldx, use_dinv)
    assert code == 0
    got = _host(dX)[:k]
    ref = sla.solve_triangular(R, X0)
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    assert rel <= 1e-12, rel
    assert np.array_equal(_host(dX)[k:], Xp[k:])  # rows past k untouched
    torch.cuda.synchronize()


def test_gram_rejects_bad_shapes(librbrp, hooks):
    import torch
    z = torch.zeros(64, dtype=torch.float64, device="cuda")
    assert hooks.rbrp_test_gram(_stream(), 129, 16, _ptr(z), 16, _ptr(z), 129, 1) == -2
    assert hooks.rbrp_test_gram(_stream(), 4, 8, _ptr(z), 16, _ptr(z), 4, 1) == -3
    assert hooks.rbrp_test_gram(_stream(), 4, 16, _ptr(z), 17, _ptr(z), 4, 1) == -5


def test_gemm_nn64_rejects_bad_shapes(librbrp, hooks):
    import torch
    z = torch.zeros(64, dtype=torch.float64, device="cuda")
    assert hooks.rbrp_test_gemm_nn64(_stream(), 6, 1, 1, _ptr(z), 6, _ptr(z), 1, _ptr(z), 6, 0) == -2
    assert hooks.rbrp_test_gemm_nn64(_stream(), 4, 1, 1, _ptr(z), 5, _ptr(z), 1, _ptr(z), 4, 0) == -6


def test_gemm_tn64_rejects_bad_shapes(librbrp, hooks):
    import torch
    z = torch.zeros(64, dtype=torch.float64, device="cuda")
    assert hooks.rbrp_test_gemm_tn64(_stream(), 1, 1, 8, _ptr(z), _ptr(z), 8, _ptr(z), 1, 1) == -4
    assert hooks.rbrp_test_gemm_tn64(_stream(), 1, 1, 16, _ptr(z), _ptr(z), 17, _ptr(z), 1, 1) == -7


@pytest.mark.parametrize("M,N,K,mode", [(64, 64, 16, 0), (130, 70, 33, 1), (5, 3, 1, 2),
                                        (256, 200, 64, 1), (64, 1, 100, 0)])
def test_gemm_nn_matches_numpy(librbrp, hooks, M, N, K, mode):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    dA, dB, dC = _cm(A), _cm(B), _cm(C0)
    ntile = (M + 63) // 64
    import torch
    dP = torch.zeros(ntile * N, dtype=torch.float64, device="cuda")
    code = hooks.rbrp_test_gemm_nn(_stream(), M, N, K, _ptr(dA), M, _ptr(dB), K, _ptr(dC), M, mode,
                                     _ptr(dP))
    assert code == 0
    AB = A @ B
    ref = AB if mode == 0 else (C0 - AB if mode == 1 else C0 + AB)
    got = _host(dC)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * np.sqrt(K + 1)
    parts = _host(dP).reshape(ntile, N)
    np.testing.assert_allclose(parts.sum(axis=0), np.sum(ref * ref, axis=0), rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("M,N,K", [(128, 32, 16), (200, 70, 33), (5, 3, 1), (1000, 130, 64),
                                   (300, 17, 128), (64, 1, 100), (5000, 300, 64), (129, 33, 97),
                                   (70000, 40, 60), (50048, 300, 64), (640, 257, 33), (64, 1, 7),
                                   (1024, 129, 64), (4096, 96, 1)])
def test_gemm_trail_update_matches_numpy(librbrp, hooks, M, N, K):
    """The trailing-update launcher against NumPy: shapes with M a multiple of 64
    and K <= 64 take the 128-column kernel with the A block shared through LDS,
    the others the 32-column kernel."""
    import torch
    rng = np.random.default_rng(M * 3 + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    dA, dC = _cm(A), _cm(C0)
    dBt = torch.as_tensor(np.ascontiguousarray(B)).cuda()   # row-major (k, c) at k*N + c
    ntile = (M + 63) // 64
    dP = torch.zeros(ntile * N, dtype=torch.float64, device="cuda")
    code = hooks.rbrp_test_gemm_upd(_stream(), M, N, K, _ptr(dA), M, _ptr(dBt), N, _ptr(dC), M,
                                      _ptr(dP))
    assert code == 0
    ref = C0 - A @ B
    got = _host(dC)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()) * np.sqrt(K + 1)
    parts = _host(dP).reshape(ntile, N)
    np.testing.assert_allclose(parts.sum(axis=0), np.sum(ref * ref, axis=0), rtol=1e-12, atol=1e-300)


# kernel 0: the library's choice for b; 1: one wave, lane = column (b <= 64);
# 2: register tiles over a workgroup (b <= 128)
_PANEL_SMALL = [(1, 8, 0), (8, 8, 0), (40, 40, 0), (64, 64, 5), (64, 64, 30)]
_PANEL_CASES = ([c + (1,) for c in _PANEL_SMALL] + [c + (2,) for c in _PANEL_SMALL]
                + [(65, 128, 0, 2), (96, 96, 5, 2), (100, 128, 10, 2), (128, 128, 0, 2),
                   (128, 128, 30, 2), (40, 64, 0, 0), (100, 128, 10, 0)])


@pytest.mark.parametrize("s,b,deps,kernel", _PANEL_CASES)
def test_panel_pivots_filter_match_dgeqp3(librbrp, hooks, s, b, deps, kernel):
    import torch
    rng = np.random.default_rng(s * 7 + deps)
    m = 600
    V = rng.standard_normal((m, s)) * np.exp(-rng.random(s) * 8)
    for q in range(deps):   # nearly dependent columns for the filter to drop
        a, c = rng.integers(0, s, size=2)
        V[:, a] = V[:, c] * 2.0 + 1e-9 * rng.standard_normal(m)
    G = V.T @ V
    i = 1000
    J = (i + rng.permutation(4 * s)[:s]).astype(np.int32)
    dG = _cm(G)
    dJ = torch.as_tensor(J).cuda()
    dJk = torch.zeros(b, dtype=torch.int32, device="cuda")
    dJin = torch.zeros(b, dtype=torch.int32, device="cuda")
    dJout = torch.zeros(b, dtype=torch.int32, device="cuda")
    dRb = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    dRinv = torch.full((b * b,), np.nan, dtype=torch.float64, device="cuda")
    bt, ns, status = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    code = hooks.rbrp_test_panel_kernel(_stream(), _ptr(dG), s, b, 0, i, _ptr(dJ), _ptr(dJk),
                                          _ptr(dRb), _ptr(dRinv), _ptr(dJin), _ptr(dJout),
                                          ctypes.byref(bt), ctypes.byref(ns),
                                          ctypes.byref(status), kernel)
    assert code == 0 and status.value == 0
    _, R, piv = sla.qr(V, mode="economic", pivoting=True)
    bt_ref = robust_filter(R, s, b, "sampled")
    assert bt.value == bt_ref
    Jk_ref = J[piv[:bt_ref]]
    assert np.array_equal(_host(dJk)[:bt_ref], Jk_ref)
    Rb = _host(dRb).reshape(b, b).T[:bt_ref, :bt_ref]
    Rref = np.abs(R[:bt_ref, :bt_ref])
    assert np.abs(np.abs(Rb) - Rref).max() <= 1e-9 * Rref.max()
    # the factor's inverse: upper triangular, and Rb Rinv = I to the conditioning of Rb
    Rinv = _host(dRinv).reshape(b, b).T[:bt_ref, :bt_ref]
    assert np.array_equal(np.tril(Rinv, -1), np.zeros_like(Rinv))
    kappa = np.linalg.cond(Rb)
    assert np.abs(Rb @ Rinv - np.eye(bt_ref)).max() <= 1e-14 * kappa * bt_ref
    np.testing.assert_allclose(Rinv, sla.solve_triangular(Rb, np.eye(bt_ref)),
                               rtol=0, atol=1e-14 * kappa * np.abs(Rinv).max() * bt_ref)
    jin, jout = symmetric_difference(i, bt_ref, Jk_ref)
    assert ns.value == len(jin)
    assert np.array_equal(_host(dJin)[:len(jin)], jin)
    assert np.array_equal(_host(dJout)[:len(jout)], jout)


_CHOL2_SMALL = [(1, 8), (5, 8), (16, 64), (17, 64), (40, 64), (63, 64), (64, 64)]
_CHOL2_CASES = ([c + (1,) for c in _CHOL2_SMALL] + [c + (2,) for c in _CHOL2_SMALL]
                + [(65, 128, 2), (80, 96, 2), (100, 128, 2), (128, 128, 2), (100, 128, 0)])


@pytest.mark.parametrize("bt,b,kernel", _CHOL2_CASES)
def test_chol2_factor_inverse_product(librbrp, hooks, bt, b, kernel):
    """Second Cholesky-QR pass: R2 = chol(G2) (upper, positive diagonal), R2inv =
    R2^{-1}, Rout = R2 Rb, against LAPACK on a near-orthonormal block's Gram."""
    import torch
    rng = np.random.default_rng(bt * 131 + b)
    Q1 = np.linalg.qr(rng.standard_normal((500, bt)))[0] + 1e-3 * rng.standard_normal((500, bt))
    G2 = Q1.T @ Q1
    Rb = np.triu(rng.standard_normal((bt, bt))) + 4.0 * np.eye(bt)
    Rbf = np.zeros((b, b))
    Rbf[:bt, :bt] = Rb
    dG2, dRb = _cm(G2), _cm(Rbf)
    dR2 = torch.full((b * b,), np.nan, dtype=torch.float64, device="cuda")
    dR2inv = torch.full((b * b,), np.nan, dtype=torch.float64, device="cuda")
    dRout = torch.full((bt * bt,), np.nan, dtype=torch.float64, device="cuda")
    status = ctypes.c_int32()
    code = hooks.rbrp_test_chol2_kernel(_stream(), _ptr(dG2), bt, b, _ptr(dRb), _ptr(dR2),
                                          _ptr(dR2inv), _ptr(dRout), ctypes.byref(status), kernel)
    assert code == 0 and status.value == 0
    R2_ref = np.linalg.cholesky(G2).T
    R2 = _host(dR2).reshape(b, b).T[:bt, :bt]
    R2inv = _host(dR2inv).reshape(b, b).T[:bt, :bt]
    Rout = _host(dRout).reshape(bt, bt).T
    np.testing.assert_allclose(R2, R2_ref, rtol=0, atol=1e-13)
    assert np.array_equal(np.tril(R2inv, -1), np.zeros((bt, bt)))
    np.testing.assert_allclose(R2inv, np.linalg.inv(R2_ref), rtol=0, atol=1e-13)
    np.testing.assert_allclose(Rout, R2_ref @ Rb, rtol=0, atol=1e-12 * np.abs(Rb).max())


@pytest.mark.parametrize("N,i,b,skew,seed", [(10, 0, 4, 1, 0), (4000, 0, 64, 1, 1),
                                             (4000, 1800, 64, 6, 2), (300, 250, 64, 1, 3),
                                             (5000, 10, 128, 12, 4), (1, 0, 8, 1, 5)])
def test_sampler_matches_oracle(librbrp, hooks, N, i, b, skew, seed):
    import torch
    rng = np.random.default_rng(seed)
    diags = rng.random(N) ** skew
    diags[rng.random(N) < 0.1] = 0.0
    if N > 1:
        diags[rng.integers(i, N)] = diags.sum() * 50   # dominant weight: redraw rounds
    kmax = N
    w = diags[i:]
    s_ref = min(b, kmax - i, int(np.count_nonzero(w > 0)))
    ga = uniform_stream(seed, 9)
    ga.random(7)   # 7 draws already consumed
    Js, d = sample_without_replacement(w, s_ref, ga)
    dd = torch.as_tensor(diags).cuda()
    dJ = torch.zeros(b, dtype=torch.int32, device="cuda")
    s_out, draws = ctypes.c_int32(), ctypes.c_uint64()
    code = hooks.rbrp_test_sample(_stream(), _ptr(dd), i, N, kmax, b, seed, 9, 7, _ptr(dJ),
                                    ctypes.byref(s_out), ctypes.byref(draws))
    assert code == 0
    assert s_out.value == s_ref
    assert np.array_equal(_host(dJ)[:s_ref], i + Js)
    assert draws.value == 7 + d


@pytest.mark.parametrize("N,i,b,skew,dominant,seed", [
    (70000, 0, 64, 1, False, 11),        # just above the single-workgroup range
    (300000, 1000, 128, 8, True, 12),    # skewed + dominant weight: several rounds
    (1 << 20, 0, 64, 1, False, 13),
    (200000, 5, 256, 20, True, 14),      # most weight on a few hundred entries
])
def test_wide_sampler_matches_oracle(librbrp, hooks, N, i, b, skew, dominant, seed):
    """The multi-workgroup sampler (candidate range > 65536) against the oracle:
    same indices in the same order, same number of uniforms consumed."""
    import torch
    rng = np.random.default_rng(seed)
    diags = rng.random(N) ** skew
    diags[rng.random(N) < 0.1] = 0.0
    if dominant:
        diags[rng.integers(i, N, size=3)] = diags.sum() * 20
    kmax = N
    w = diags[i:]
    s_ref = min(b, kmax - i, int(np.count_nonzero(w > 0)))
    ga = uniform_stream(seed, 9)
    ga.random(3)
    Js, d = sample_without_replacement(w, s_ref, ga)
    dd = torch.as_tensor(diags).cuda()
    dJ = torch.zeros(b, dtype=torch.int32, device="cuda")
    s_out, draws = ctypes.c_int32(), ctypes.c_uint64()
    code = hooks.rbrp_test_sample(_stream(), _ptr(dd), i, N, kmax, b, seed, 9, 3, _ptr(dJ),
                                  ctypes.byref(s_out), ctypes.byref(draws))
    assert code == 0
    assert s_out.value == s_ref
    assert np.array_equal(_host(dJ)[:s_ref], i + Js)
    assert draws.value == 3 + d
    print(f"[wide sampler N={N}] s={s_ref} rounds' draws={d}")
