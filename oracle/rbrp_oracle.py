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
  ``||A(:,rd) - A(:,J) X||_F^2 == sum(residual diags)`` holds.
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


def sample_without_replacement(w, s: int, rng: np.random.Generator):
    """Weighted sampling of ``s`` distinct indices of ``w`` (rounds algorithm).

    Restates ``rng.choice(len(w), size=s, p=w/sum(w), replace=False)`` as used
    at RBRP.ipynb c11:108: each round draws ``s - n_found`` uniforms, zeroes
    the weights already found, forms ``cdf = cumsum(w) / cdf[-1]``, takes
    ``searchsorted(cdf, u, side='right')`` and appends the new distinct picks in
    order of first occurrence.  Returns ``(indices, draws_consumed)``.
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
    blocks: list              # per-block dict(i, s, bt, J, Jk)
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
        Js, d = sample_without_replacement(w, s, rng)            # c11:108
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
        blocks.append(dict(i=i, s=s, bt=bt, J=J.copy(), Jk=Jk.copy()))
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
