"""CPU restatement of the sketch-based ID comparators (oracle; test infrastructure only).

``rand_cpqr_os`` restates gpu/randCPQR_OS.cu:20-111 (spec matlab/randCPQR_OS.m:1-49)
and ``rand_lupp_os`` restates gpu/randLUPP_OS.cu:14-119 (spec
matlab/randLUPP_OS.m:1-54).  ``rand_adap_lupp`` restates blocked_rlupp_id
(python/cache.ipynb c8:80-143), the only definition of the behaviour that
gpu/rid.hpp:5-7 declares as ``RandAdapLUPP``.  Both take the Gaussian sketch matrix ``G`` as an
explicit argument so that a device implementation and this oracle can be fed
the identical sketch.  Pivoted QR and LU come from SciPy's LAPACK
(``dgeqp3``, ``dgetrf``), which is what the reference calls through MAGMA and
cuSOLVER.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def pivots_to_permutation(ipiv, P, ofs: int = 0):
    """Apply LAPACK row interchanges ``ipiv`` (0-based here) to ``P`` in order.

    gpu/permute.cu:8-23 (the reference's ``ipiv`` is 1-based and subtracts 1).
    """
    P = np.array(P, dtype=np.int64, copy=True)
    for j, pj in enumerate(np.asarray(ipiv, dtype=np.int64)):
        a, c = j + ofs, int(pj) + ofs
        P[a], P[c] = P[c], P[a]
    return P


def compute_interpolation_matrix_os(Y, P, k: int):
    """Oversampled interpolation matrix ``T`` (k x (n-k)) from sketch ``Y`` (r x n).

    gpu/interp.cu:45-105: ``Yp = Y(:,P)``; QR of ``Yp(:,1:k)`` (geqrf :82);
    ``Q^T`` applied to ``Yp(:,k+1:n)`` (ormqr :87); triangular solve with the
    ``k x k`` R (trsm :95); ``T`` is the top ``k`` rows.
    """
    Y = np.asarray(Y, dtype=np.float64)
    Yp = Y[:, np.asarray(P, dtype=np.int64)]
    Q1, R1 = sla.qr(Yp[:, :k], mode="economic")
    return sla.solve_triangular(R1, Q1.T @ Yp[:, k:], lower=False)


def rand_cpqr_os(A, r: int, G):
    """Randomised CPQR ID with 2x oversampling; returns ``(sk, rd, T)``.

    gpu/randCPQR_OS.cu: ``Y = G A`` (:42, G is 2r x m), pivoted QR of
    ``Y(1:r,:)`` (:64), ``sk = jpvt[:r]``, ``rd = jpvt[r:]`` (:91-92), ``T`` by
    the oversampled interpolation (:95).
    """
    A = np.asarray(A, dtype=np.float64)
    Y = np.asarray(G, dtype=np.float64) @ A
    _, _, jpvt = sla.qr(Y[:r, :], mode="economic", pivoting=True)
    T = compute_interpolation_matrix_os(Y, jpvt, r)
    return jpvt[:r].copy(), jpvt[r:].copy(), T


def rand_lupp_os(A, k: int, G):
    """Randomised LUPP ID with 2x oversampling; returns ``(sk, rd, T)``.

    gpu/randLUPP_OS.cu: ``Y = G A`` (:43, G is 2k x m), LU with partial
    pivoting of ``Y(1:k,:)^T`` (n x k, :57-81), permutation from the pivots
    (:85, gpu/permute.cu:8-23), ``T`` by the oversampled interpolation (:96).
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[1]
    Y = np.asarray(G, dtype=np.float64) @ A
    _, ipiv = sla.lu_factor(Y[:k, :].T.copy(), check_finite=False)
    P = pivots_to_permutation(ipiv, np.arange(n))
    T = compute_interpolation_matrix_os(Y, P, k)
    return P[:k].copy(), P[k:].copy(), T


def rand_cpqr(A, r: int, G):
    """Randomised CPQR ID without oversampling; returns ``(sk, rd, T)``.

    gpu/randCPQR.cu (spec matlab/randCPQR.m): ``Y = G A`` (:42, G is r x m),
    pivoted QR of all of ``Y`` (:52), ``sk = jpvt[:r]``, ``rd = jpvt[r:]``
    (:70-72), ``T = R11^{-1} R12`` from that factor (:60-64).
    """
    A = np.asarray(A, dtype=np.float64)
    Y = np.asarray(G, dtype=np.float64) @ A
    _, R, jpvt = sla.qr(Y, mode="economic", pivoting=True)
    T = sla.solve_triangular(R[:r, :r], R[:r, r:], lower=False)
    return jpvt[:r].copy(), jpvt[r:].copy(), T


def rand_lupp(A, k: int, G):
    """Randomised LUPP ID without oversampling; returns ``(sk, rd, T)``.

    gpu/randLUPP.cu (spec matlab/randLUPP.m): ``Y = G A`` (:43, G is k x m), LU
    with partial pivoting of ``Y^T`` (:47-78), permutation from the pivots
    (:82-85), ``T = (L21 L11^{-1})^T`` (:89-106).
    """
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[1]
    Y = np.asarray(G, dtype=np.float64) @ A
    LU, ipiv = sla.lu_factor(Y.T.copy(), check_finite=False)
    P = pivots_to_permutation(ipiv, np.arange(n))
    L11 = np.tril(LU[:k, :k], -1) + np.eye(k)
    L21 = LU[k:, :k]
    T = sla.solve_triangular(L11, L21.T, lower=True, trans="T", unit_diagonal=True)
    return P[:k].copy(), P[k:].copy(), T


def rand_adap_lupp(A, tol: float, b: int, sketch):
    """Adaptive randomised LUPP ID; returns ``(sk, rd, T, rank, est)``.

    Column form of blocked_rlupp_id (python/cache.ipynb c8:80-143) with one
    estimation sketch per block (its ``num_smp`` = 1, as
    blocked_rlupp_id_local, c8:1-60).  ``sketch(t)`` returns block t's sketch
    matrix ``G_t`` (b x m).  Block 0's residual is ``(G_0 A)^T``; each block
    factors the candidates' residual (candidates x bt) with partial pivoting
    (c8:94), moves its first bt pivots to the skeleton (c8:106-107), and keeps
    the interpolation matrix ``M = L2 L1^{-1}`` of the accumulated unit-lower
    factor (c8:97-103, 111-116) up to date; the next sketch ``Y`` gives the
    next residual ``Y(:, cand)^T - M Y(:, skel)^T`` (c8:115-116) and the error
    estimate, its Frobenius norm (c8:117).  Stops when the estimate is at most
    ``tol * ||A||_F`` (the notebook compares with an absolute tol, c8:128) or
    the rank reaches min(m, n).  ``T = M^T``, so ``A(:, rd) ~ A(:, sk) T``;
    ``est`` is the last estimate relative to ``||A||_F``.
    """
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    p = min(m, n)
    normA2 = float(np.sum(A * A))
    P = np.arange(n)
    M = np.zeros((n, p))
    Rt = np.zeros((n, b))
    Y = np.asarray(sketch(0), dtype=np.float64) @ A
    Rt[:, :] = Y[:, P].T
    k, t = 0, 0
    est2 = normA2
    while k < p:
        bt = min(b, p - k)           # (never past min(m, n))
        nc = n - k
        lu, ipiv = sla.lu_factor(Rt[k:, :bt].copy(), check_finite=False)
        L = np.tril(lu, -1)[:, :bt] + np.eye(nc, bt)
        for j, q in enumerate(np.asarray(ipiv, dtype=np.int64)):
            if q != j:
                a, c = k + j, k + int(q)
                M[[a, c], :k] = M[[c, a], :k]
                P[[a, c]] = P[[c, a]]
        if nc > bt:
            Mb = sla.solve_triangular(L[:bt].T, L[bt:].T, lower=False, unit_diagonal=True).T
            if k > 0:
                M[k + bt:, :k] -= Mb @ M[k:k + bt, :k]
            M[k + bt:, k:k + bt] = Mb
        k += bt
        t += 1
        if k >= p:
            est2 = 0.0
            break
        Y = np.asarray(sketch(t), dtype=np.float64) @ A
        Rt[k:, :] = Y[:, P[k:]].T - M[k:, :k] @ Y[:, P[:k]].T
        est2 = float(np.sum(Rt[k:, :] ** 2))
        if est2 <= tol * tol * normA2:
            break
    T = M[k:, :k].T.copy()
    est = float(np.sqrt(est2 / normA2)) if normA2 > 0 else 0.0
    return P[:k].copy(), P[k:].copy(), T, k, est
