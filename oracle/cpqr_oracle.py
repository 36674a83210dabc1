"""CPU restatement of the full-matrix CPQR ID baseline (oracle; test infrastructure only).

``cpqr_id`` restates ``CPQR`` of gpu/id_cpqr.cu:30-172: a column-pivoted QR of
all of ``A`` (the reference calls MAGMA ``dgeqp3``; here SciPy's LAPACK
``dgeqp3``), then for each requested rank r the interpolation matrix
``T = R11^{-1} R12`` (``cublasDtrsm`` at :133-136) and the squared relative
error ``||A(:, P(r:n)) - A(:, P(0:r)) T||_F^2 / ||A||_F^2`` (:142-166).
``partial_time_fraction`` restates the partial-factorization time model of
:93-104.

Pins (tests/test_cpqr.py): the pivots are LAPACK's; each error equals
``||R22||_F^2 / ||A||_F^2`` read off the triangular factor (the exact identity
of a QR-based ID, computed without forming T); the errors are no smaller than
the SVD optimum at the same rank; the time fraction equals a brute-force sum
of per-step work and is 1 at full rank.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def partial_time_fraction(m: int, n: int, r: int) -> float:
    """Share of a full pivoted-QR time charged to its first ``r`` steps (gpu/id_cpqr.cu:93-104).

    Step j (0-based) is charged (m - j)(n - j); with q = min(m, n) and
    p = max(m, n) - q the charge of the step with t = q - j columns left is
    (p + t) t, so the first r steps cost S(q) - S(q - r) with
    S(t) = p t (t + 1) / 2 + t (t + 1)(2t + 1) / 6.
    """
    q = min(m, n)
    p = max(m, n) - q

    def S(t):
        return p * t * (t + 1) / 2.0 + t * (t + 1) * (2 * t + 1) / 6.0

    return (S(q) - S(q - r)) / S(q)


def flops_dgeqrf(m: float, n: float) -> float:
    """LAPACK Working Note 41 Householder QR flop count (gpu/flops.hpp:31-40)."""
    if m > n:
        muls = n * (n * (0.5 - n / 3.0 + m) + m + 23.0 / 6.0)
        adds = n * (n * (0.5 - n / 3.0 + m) + 5.0 / 6.0)
    else:
        muls = m * (m * (-0.5 - m / 3.0 + n) + 2.0 * n + 23.0 / 6.0)
        adds = m * (m * (-0.5 - m / 3.0 + n) + n + 5.0 / 6.0)
    return muls + adds


def cpqr_id(A, ranks):
    """Full-matrix CPQR ID errors at ``ranks`` (gpu/id_cpqr.cu:30-172).

    Returns ``(errors, piv, R, flops)``: squared relative errors per rank, the
    0-based pivot order, the triangular factor and the reference's flop count
    FLOPS_DGEQRF(m, n) + sum r^2 (n - r) (:70, :139).
    """
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    anrm2 = float(np.sum(A * A))                                   # :37-38
    _, R, piv = sla.qr(A, mode="economic", pivoting=True)           # :62
    flops = flops_dgeqrf(m, n)                                      # :70
    errs = []
    for r in ranks:
        if r >= n:
            errs.append(0.0)
            continue
        T = sla.solve_triangular(R[:r, :r], R[:r, r:])              # :133-136
        flops += float(r) * r * (n - r)                             # :139
        E = A[:, piv[r:]] - A[:, piv[:r]] @ T                       # :142-159
        errs.append(float(np.sum(E * E)) / anrm2)                   # :162-166
    return errs, piv, R, flops
