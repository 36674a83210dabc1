"""Full-matrix CPQR ID baseline (gpu/id_cpqr.cu:30-172): oracle pins (CPU) and the HIP library (GPU)."""

import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import cpqr_id, partial_time_fraction
from oracle.cpqr_oracle import flops_dgeqrf
from robust_blockwise_random_pivoting_amd.matrices import exp_spectrum, spectrum_matrix


def _mat(m, n, rate, seed):
    return spectrum_matrix(m, n, exp_spectrum(min(m, n), rate), seed)


# ---------------------------------------------------------------- CPU: the oracle

def test_time_fraction_is_a_sum_of_step_work():
    for m, n in [(50, 30), (30, 50), (40, 40), (1000, 7)]:
        q = min(m, n)
        work = np.array([(m - j) * (n - j) for j in range(q)], dtype=np.float64)
        for r in range(q + 1):
            assert partial_time_fraction(m, n, r) == pytest.approx(work[:r].sum() / work.sum(),
                                                                   rel=1e-12, abs=1e-15)
        assert partial_time_fraction(m, n, q) == pytest.approx(1.0)


def test_flop_count_matches_lawn41_leading_terms():
    # 2 m n^2 - 2 n^3 / 3 for m >= n and 2 n m^2 - 2 m^3 / 3 otherwise, to leading order
    for m, n in [(5000, 400), (400, 5000)]:
        a, b = max(m, n), min(m, n)
        assert flops_dgeqrf(m, n) == pytest.approx(2.0 * a * b * b - 2.0 * b ** 3 / 3.0, rel=5e-3)


@pytest.mark.parametrize("m,n,rate,seed", [(200, 120, 30.0, 1), (90, 150, 20.0, 2), (150, 150, 1e9, 3)])
def test_oracle_error_is_the_trailing_triangle_and_above_svd_optimum(m, n, rate, seed):
    A = _mat(m, n, rate, seed) if rate < 1e8 else np.random.default_rng(seed).standard_normal((m, n))
    q = min(m, n)
    ranks = [1, q // 4, q // 2, q - 1, q]
    errs, piv, R, fl = cpqr_id(A, ranks)
    _, R2, piv2 = sla.qr(A, mode="economic", pivoting=True)
    assert np.array_equal(piv, piv2)
    anrm2 = np.sum(A * A)
    sv = np.linalg.svd(A, compute_uv=False)
    for r, e in zip(ranks, errs):
        r22 = np.sum(R[r:, r:] ** 2) / anrm2          # ||A(:,rd) - A(:,sk)T||^2 = ||R22||^2
        assert e == pytest.approx(r22, rel=1e-8, abs=1e-14)
        assert e >= np.sum(sv[r:] ** 2) / anrm2 * (1 - 1e-10) - 1e-15
    assert errs[-1] == 0.0 or errs[-1] <= 1e-25
    assert all(x >= y for x, y in zip(errs, errs[1:]))
    assert fl == pytest.approx(flops_dgeqrf(m, n) + sum(r * r * (n - r) for r in ranks if r < n))


def test_capi_rejects_a_null_handle(librbrp):
    err = np.zeros(1)
    rk = np.array([1], dtype=np.int32)
    tq, fl = ctypes.c_double(), ctypes.c_double()
    code = librbrp.rbrp_cpqr_id_d(None, None, 4, 4, 4, 1, ctypes.c_void_p(rk.ctypes.data),
                                  ctypes.c_void_p(err.ctypes.data), ctypes.c_void_p(err.ctypes.data),
                                  ctypes.byref(tq), ctypes.byref(fl), None)
    assert code == -1


# ---------------------------------------------------------------- GPU: the HIP library

def _dev(A):
    import torch
    return torch.as_tensor(np.ascontiguousarray(A.T)).cuda().t()


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,rate,seed", [(240, 160, 40.0, 11), (120, 200, 30.0, 12),
                                           (300, 300, 1e9, 13), (1000, 64, 10.0, 14),
                                           (9000, 48, 12.0, 15), (20000, 96, 25.0, 16),
                                           # many rejected norm downdates: panels cut short
                                           # (tall) and norms recomputed in the step (short)
                                           (6000, 300, 20.0, 17), (400, 1200, 50.0, 18)])
def test_cpqr_matches_oracle(librbrp, m, n, rate, seed):
    from robust_blockwise_random_pivoting_amd import CPQR
    A = _mat(m, n, rate, seed) if rate < 1e8 else np.random.default_rng(seed).standard_normal((m, n))
    q = min(m, n)
    ranks = sorted({1, q // 8 + 1, q // 2, q - 1, q})
    print(f"cpqr case {m}x{n} ranks {ranks}", flush=True)
    err, tr, tq, fl, perm = CPQR(_dev(A), ranks)
    e_ref, piv, R, fl_ref = cpqr_id(A, ranks)
    assert np.array_equal(perm.cpu().numpy(), piv)
    for r, e, er in zip(ranks, err, e_ref):
        assert e == pytest.approx(er, rel=1e-7, abs=1e-14), r
    assert fl == pytest.approx(fl_ref, rel=1e-12)
    assert tq > 0
    # t_ranks = the reference's time model plus the solve time (none at r = n)
    for r, t in zip(ranks, tr):
        assert t >= tq * partial_time_fraction(m, n, r) * (1 - 1e-9)
        if r == n:
            assert t == pytest.approx(tq, rel=1e-12)


@pytest.mark.gpu
def test_cpqr_argument_errors(librbrp):
    from robust_blockwise_random_pivoting_amd import Handle
    import torch
    A = _dev(np.random.default_rng(0).standard_normal((20, 10)))
    h = Handle()
    err, trs = np.zeros(2), np.zeros(2)
    tq, fl = ctypes.c_double(), ctypes.c_double()

    def call(ranks, m=20, n=10, lda=20, nr=None):
        rk = np.asarray(ranks, dtype=np.int32)
        return librbrp.rbrp_cpqr_id_d(h.ptr, ctypes.c_void_p(A.data_ptr()), m, n, lda,
                                      len(rk) if nr is None else nr, ctypes.c_void_p(rk.ctypes.data),
                                      ctypes.c_void_p(err.ctypes.data), ctypes.c_void_p(trs.ctypes.data),
                                      ctypes.byref(tq), ctypes.byref(fl), None)
    assert call([0]) == -7
    assert call([11]) == -7
    assert call([3], nr=-1) == -6
    assert call([3], lda=19) == -5
    assert call([3], m=0) == -3
    assert call([3, 10]) == 0 and err[1] == 0.0
    torch.cuda.synchronize()
    h.close()
