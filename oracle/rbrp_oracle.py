"""CPU restatement of the RBRP column interpolative decomposition (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker / timed CPU
baseline.  Never imported by the product package.

Semantics (column ID of a dense FP64 ``A`` of shape m x n, column-major):

1. ``diags[j] = ||A(:,j)||^2`` and ``N0 = sum(diags)``       gpu/rbrp.cu:352-355
2. loop while ``i < p = min(m, n)`` (or ``max_rank``)         gpu/rbrp.cu:367
   a. ``s = min(b, p - i, nnz(diags[i:]))``                  matlab/RBRP.m:32,
                                                              RBRP.ipynb c11:106
   b. sample ``s`` distinct positions of ``[i, n)`` with weights ``diags[i:]``
      by NumPy's weighted-without-replacement rounds algorithm fed by a
      Philox4x64-10 uniform stream                            RBRP.ipynb c11:108
   c. re-orthogonalise the sampled residual columns against the skeleton
      basis once more (second Gram-Schmidt pass; the trailing matrix already
      holds the first pass)                                   matlab/RBRP.m:55
   d. pivoted Householder QR of the sampled block (LAPACK ``dgeqp3``)
                                                              gpu/rbrp.cu:390-417
   e. robust filter on the block's R                          gpu/rbrp.cu:149-208
   f. symmetric-difference column permutation (ascending pairing)
                                                              gpu/rbrp.cu:211-294
   g. rank-b update of the trailing residual, ``R(blk, rem) = Q_b^T A_rem``
                                                              gpu/rbrp.cu:297-330
   h. residual column norms recomputed exactly from the updated residual
      (the reference downdates, gpu/rbrp.cu:455-465; the values are equal in
      exact arithmetic and the recompute has no cancellation floor)
   i. ``err_t = sum(diags[rem]) / N0`` (squared)              gpu/rbrp.cu:467
   j. stop when ``err_t <= tol**2`` (``tol`` is the non-squared relative
      Frobenius tolerance)                                    RBRP.ipynb c11:164
3. ``X = R11^{-1} R12``                                        matlab/driver.m:86

Pins (each checked by ``tests/test_oracle.py``):

* uniform stream: ``numpy.random.Philox`` against the Random123 known-answer
  vectors (Salmon et al., SC'11);
* sampler: ``numpy.random.Generator.choice(replace=False, p=...)`` on the same
  stream (identical indices and identical final generator state);
* pivoted QR: SciPy's LAPACK ``dgeqp3`` is called directly; invariants
  ``A P = Q R`` and non-increasing ``|R_jj|``;
* filter: brute force on tiny R and the ``>=`` boundary case;
* permutation: brute-force check of the reference's three copy statements;
* whole loop: an independent left-looking restatement of matlab/RBRP.m fed the
  same uniforms selects the same skeleton, and the QR identity
  ``||A(:,rd) - A(:,J) X||_F^2 == sum(residual diags)`` holds;
* near-tie margins (``block_margins``, ``first_near_tie``): closed forms on
  planted ties (a uniform at a known distance from a CDF step, two columns of
  known norms, a filter sum at a known distance from its threshold).

Near ties.  Each block records how close its discrete decisions came to
flipping: the sampler's CDF margin (the distance of every uniform from the
CDF steps around the index it selected, in units of the total weight), the
pivot margin (the relative gap between the chosen column's partial norm and
the largest other one at every pivoted-QR step), the filter margin (the
relative distance of the filter's running sums from its threshold) and the
stop margin (the relative distance of the block's squared error from tol^2).
An implementation that forms the same quantities in another order or by
another factorization may decide differently only where a margin is below
its rounding error; ``first_near_tie`` names the first such block.
"""

from __future__ import annotations

import dataclasses
from typing import List

import numpy as np
import scipy.linalg as sla
from scipy.linalg import blas as sblas


def uniform_stream(seed: int, offset: int = 0) -> np.random.Generator:
    """Philox4x64-10 uniform stream with key ``(seed, 0)`` and counter ``offset``.

    Replaces the reference's ``thrust::default_random_engine`` seeded 1986 with
    a global skip counter (gpu/random.cu:26-54); the Python reference draws
    from ``default_rng(42)`` (RBRP.ipynb c11:89).  Draw ``t`` (0-based) is word
    ``t % 4`` of Philox(counter = offset + 1 + t // 4, key), converted to a
    double as ``(x >> 11) * 2**-53`` (NumPy's ``random()``).
    """
    return np.random.Generator(np.random.Philox(key=int(seed), counter=int(offset)))


def sample_without_replacement(w, s: int, rng: np.random.Generator, margins=None):
    """Weighted sampling of ``s`` distinct indices of ``w`` (rounds algorithm).

    Restates ``rng.choice(len(w), size=s, p=w/sum(w), replace=False)`` as used
    at RBRP.ipynb c11:108: each round draws ``s - n_found`` uniforms, zeroes
    the weights already found, forms ``cdf = cumsum(w) / cdf[-1]``, takes
    ``searchsorted(cdf, u, side='right')`` and appends the new distinct picks in
    order of first occurrence.  Returns ``(indices, draws_consumed)``.  With
    ``margins`` (a list), appends each uniform's distance from the nearer of
    the two CDF values around its index, ``min(cdf[j] - u, u - cdf[j-1])``.
    """
    w = np.array(w, dtype=np.float64, copy=True)
    if s > np.count_nonzero(w > 0):
        raise ValueError("fewer non-zero weights than samples")
    found: List[int] = []
    draws = 0
    while len(found) < s:
        r = s - len(found)
        u = rng.random(r)
        draws += r
        if found:
            w[np.asarray(found)] = 0.0
        cdf = np.cumsum(w)
        cdf /= cdf[-1]
        new = cdf.searchsorted(u, side="right")
        if margins is not None:
            lower = np.where(new > 0, cdf[np.maximum(new - 1, 0)], 0.0)
            margins.extend(np.minimum(cdf[new] - u, u - lower).tolist())
        _, first = np.unique(new, return_index=True)
        first.sort()
        found.extend(int(x) for x in new[first])
    return np.asarray(found, dtype=np.int64), draws


def robust_filter(R, s: int, b: int, mode: str = "sampled") -> int:
    """Number of pivoted columns kept from an ``s x s`` upper-triangular ``R``.

    ``sqr[r] = sum_{c>=r} R[r,c]^2``; ``csum`` = running sum of ``sqr`` from the
    last row upwards; keep ``s - k`` where ``k`` is the first index with
    ``csum[k] >= csum[-1] / s`` (gpu/rbrp.cu:193-207, mode ``"sampled"``).
    Mode ``"nominal"`` divides by the nominal block size ``b`` instead, which is
    the Python reference's ``tol_loc = 1/b`` rule (RBRP.ipynb c11:69-73,90).
    """
    R = np.asarray(R, dtype=np.float64)
    sqr = np.zeros(s)
    for r in range(s):
        acc = 0.0
        for c in range(r, s):
            acc += R[r, c] * R[r, c]
        sqr[r] = acc
    csum = np.zeros(s)
    acc = 0.0
    for t in range(s):
        acc += sqr[s - 1 - t]
        csum[t] = acc
    div = s if mode == "sampled" else b
    thr = csum[-1] / div
    k = 0
    while k < s and not (csum[k] >= thr):
        k += 1
    return s - k


EPS = float(np.finfo(np.float64).eps)
# thresholds below which a margin is a near tie: the sampler's CDF is summed
# in another order on the device (about 1e3 eps covers its rounding); the
# device's pivots and filter sums come from a Gram matrix and a pivoted
# Cholesky factorization, whose Schur diagonals carry about s^3 eps relative
# error for a block of s columns; the error trace is compared to 1e-6.
TIE_CDF = 1e3 * EPS
TIE_STOP = 1e-6


def tie_pivot(s: int) -> float:
    """Near-tie threshold of the pivot and filter margins of an s-column block."""
    return max(1e3 * EPS, float(s) ** 3 * EPS)


def pivot_margin(Rv, steps: int = 0) -> float:
    """Smallest relative gap between the chosen column's partial norm and the
    largest other candidate's over the first ``steps`` steps (all when 0) of a
    pivoted QR with factor ``Rv`` (columns in pivot order): at step k the
    candidates t >= k have partial norms ``sqrt(sum_{r >= k} Rv[r, t]^2)`` and
    the pivot is t = k.  Steps whose chosen norm is zero are skipped; +inf when
    no step has two candidates.  (The block keeps its first bt pivots: later
    steps reach the decisions only through the filter's sums.)
    """
    Rv = np.asarray(Rv, dtype=np.float64)
    rows, s = Rv.shape
    if s < 2:
        return float("inf")
    sq = Rv * Rv
    S = np.cumsum(sq[::-1, :], axis=0)[::-1, :]       # S[k, t] = sum_{r >= k} Rv[r, t]^2
    best = float("inf")
    last = min(rows, s - 1) if steps <= 0 else min(rows, s - 1, steps)
    for k in range(last):
        chosen = np.sqrt(S[k, k])
        if chosen == 0.0:
            continue
        other = np.sqrt(np.max(S[k, k + 1:]))
        best = min(best, (chosen - other) / chosen)
    return best


def filter_margin(R, s: int, b: int, mode: str = "sampled") -> float:
    """Relative distance of the filter's deciding running sums from its
    threshold: ``min(csum[k] - thr, thr - csum[k-1]) / thr`` at the first k with
    ``csum[k] >= thr`` (see ``robust_filter``); +inf for a zero threshold.  The
    last sum is the total, at or above the threshold whatever the rounding, so
    it counts only as the lower side.
    """
    R = np.asarray(R, dtype=np.float64)
    sqr = np.array([float(np.sum(R[r, r:s] ** 2)) for r in range(s)])
    csum = np.cumsum(sqr[::-1])
    div = s if mode == "sampled" else b
    thr = csum[-1] / div
    if thr <= 0.0:
        return float("inf")
    k = int(np.argmax(csum >= thr))
    m = csum[k] - thr if k < s - 1 else float("inf")
    if k > 0:
        m = min(m, thr - csum[k - 1])
    return float(m / thr)


def first_near_tie(res):
    """``(block index, which margin, value, threshold)`` of the first block of
    an ``RBRPResult`` with a margin below its threshold, or None."""
    for t, blk in enumerate(res.blocks):
        for key, thr in (("cdf_margin", TIE_CDF), ("pivot_margin", tie_pivot(blk["s"])),
                         ("filter_margin", tie_pivot(blk["s"])), ("stop_margin", TIE_STOP)):
            if blk[key] < thr:
                return t, key, blk[key], thr
    return None


def margin_summary(res) -> str:
    """One line: each margin's minimum over the blocks and where it occurs."""
    parts = []
    for key in ("cdf_margin", "pivot_margin", "filter_margin", "stop_margin"):
        vals = [blk[key] for blk in res.blocks]
        if not vals:
            continue
        t = int(np.argmin(vals))
        parts.append(f"{key.split('_')[0]} {vals[t]:.2e} (block {t})")
    tie = first_near_tie(res)
    parts.append("near tie: none" if tie is None else
                 f"near tie: block {tie[0]} {tie[1]} {tie[2]:.2e} < {tie[3]:.2e}")
    return ", ".join(parts)


def symmetric_difference(i: int, bt: int, jk):
    """``(Jin, Jout)``: ``Jin = sort(jk) \\ [i, i+bt)``, ``Jout = [i, i+bt) \\ jk``.

    Both ascending (thrust::set_difference on sorted inputs,
    gpu/rbrp.cu:211-234).
    """
    js = np.sort(np.asarray(jk, dtype=np.int64))
    block = np.arange(i, i + bt, dtype=np.int64)
    jin = np.setdiff1d(js, block, assume_unique=True)
    jout = np.setdiff1d(block, js, assume_unique=True)
    return jin, jout


@dataclasses.dataclass
class RBRPResult:
    k: int
    perm: np.ndarray          # (n,) skeleton = perm[:k]
    X: np.ndarray             # (k, n-k) interpolation matrix R11^{-1} R12
    R: np.ndarray             # (k, n) upper-trapezoidal factor, columns in perm order
    ranks: np.ndarray         # cumulative rank after each block
    error: np.ndarray         # squared relative residual after each block
    draws: int                # uniforms consumed
    blocks: list              # per-block dict(i, s, bt, J, Jk and the four margins)
    resid: np.ndarray         # final residual column norms^2 (positions k..n-1)
    N0: float


def _colnorm2(X):
    return np.einsum("ij,ij->j", X, X)


def rbrp_colid(A, b: int, tol: float = 0.0, seed: int = 0, offset: int = 0,
               max_rank: int = 0, filter_mode: str = "sampled",
               want_x: bool = True) -> RBRPResult:
    """RBRP column ID of ``A`` (m x n FP64), right-looking form.

    Control flow follows ``RBRP`` (gpu/rbrp.cu:333-503) with the tolerance stop
    of ``blocked_rpid`` (RBRP.ipynb c11:164) and the output ``X = R11\\R12`` of
    matlab/driver.m:86.  See the module docstring for the step-by-step map.
    """
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    p = min(m, n)
    kmax = p if not max_rank else min(int(max_rank), p)
    Aw = np.array(A, order="F", copy=True)                      # rbrp.cu:337-338
    diags = _colnorm2(Aw)                                        # rbrp.cu:352-354
    N0 = float(np.sum(diags))                                    # rbrp.cu:355
    perm = np.arange(n, dtype=np.int64)
    R = np.zeros((max(kmax, 1), n), order="F")
    rng = uniform_stream(seed, offset)
    draws = 0
    ranks, errors, blocks = [], [], []
    i = 0
    tol2 = float(tol) ** 2
    while i < kmax:
        w = diags[i:]
        nnz = int(np.count_nonzero(w > 0))
        if nnz == 0:                                             # RBRP.m:113-116
            break
        s = min(b, kmax - i, nnz)                                # RBRP.m:32
        cm = []
        Js, d = sample_without_replacement(w, s, rng, cm)        # c11:108
        draws += d
        J = i + Js
        V = np.asfortranarray(Aw[:, J])
        if i > 0:                                                # RBRP.m:55
            Qp = Aw[:, :i]
            C = Qp.T @ V
            V -= Qp @ C
            R[:i, J] += C
            Aw[:, J] = V
        Qv, Rv, piv = sla.qr(V, mode="economic", pivoting=True)   # rbrp.cu:410
        bt = robust_filter(Rv, s, b, filter_mode)                # rbrp.cu:424
        Jk = J[piv[:bt]]                                         # rbrp.cu:428-434
        jin, jout = symmetric_difference(i, bt, Jk)              # rbrp.cu:239-242
        tmp = perm[Jk].copy()                                    # rbrp.cu:271-280
        perm[jin] = perm[jout]
        perm[i:i + bt] = tmp
        if i > 0:                                                # RBRP.m:91-93
            tmpR = R[:i, Jk].copy()
            R[:i, jin] = R[:i, jout]
            R[:i, i:i + bt] = tmpR
        Aw[:, jin] = Aw[:, jout]                                 # rbrp.cu:257-262
        diags[jin] = diags[jout]                                 # rbrp.cu:284-289
        Qb = np.asfortranarray(Qv[:, :bt])
        Aw[:, i:i + bt] = Qb                                     # RBRP.m:77
        R[i:i + bt, i:i + bt] = Rv[:bt, :bt]                     # RBRP.m:78
        if i + bt < n:                                           # rbrp.cu:297-330
            rem = Aw[:, i + bt:]
            W = sblas.dgemm(1.0, Qb, rem, trans_a=True)
            R[i:i + bt, i + bt:] = W
            out = sblas.dgemm(-1.0, Qb, W, beta=1.0, c=rem, overwrite_c=True)
            if not np.shares_memory(out, Aw):
                Aw[:, i + bt:] = out
            diags[i + bt:] = _colnorm2(Aw[:, i + bt:])           # rbrp.cu:455-465
        err = float(np.sum(diags[i + bt:])) / N0 if N0 > 0 else 0.0   # rbrp.cu:467
        blocks.append(dict(i=i, s=s, bt=bt, J=J.copy(), Jk=Jk.copy(),
                           cdf_margin=float(min(cm)), pivot_margin=pivot_margin(Rv, bt),
                           filter_margin=filter_margin(Rv, s, b, filter_mode),
                           stop_margin=abs(err - tol2) / tol2 if tol2 > 0 else float("inf")))
        i += bt
        ranks.append(i)
        errors.append(err)
        if err <= tol2:                                          # c11:164
            break
    k = i
    X = interpolation_matrix(R, k) if want_x else None
    return RBRPResult(k=k, perm=perm, X=X, R=R[:k].copy(), ranks=np.asarray(ranks),
                      error=np.asarray(errors), draws=draws, blocks=blocks,
                      resid=diags[k:].copy(), N0=N0)


def interpolation_matrix(R, k: int):
    """``X = R(:k,:k)^{-1} R(:k,k:)`` by an upper-triangular solve (matlab/driver.m:86)."""
    if k == 0:
        return np.zeros((0, R.shape[1]))
    return sla.solve_triangular(R[:k, :k], R[:k, k:], lower=False)


def id_error(A, perm, k: int, X):
    """Non-squared relative ID error ``||A(:,rd) - A(:,sk) X||_F / ||A||_F``.

    matlab/driver.m:88-93 (the driver squares it at :94; RBRP.ipynb c7:40-53
    reports the squared, unnormalised value).
    """
    A = np.asarray(A, dtype=np.float64)
    sk, rd = perm[:k], perm[k:]
    E = A[:, rd] - A[:, sk] @ X
    return float(np.linalg.norm(E) / np.linalg.norm(A))
