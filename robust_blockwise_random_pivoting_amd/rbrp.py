"""Host-side mirror of the reference's RBRP call surface over the HIP C ABI.

``RBRP(m, n, A, blk)`` mirrors ``void RBRP(int m, int n, double *A, int blk,
ranks, error, t_ranks, t_rbrp, flops)`` (gpu/rbrp.hpp:6-8): no tolerance, runs
to ``min(m, n)``, returns the cumulative ranks, squared relative residuals and
cumulative times per block.  ``rbrp_colid`` is the added form with a
tolerance and the skeleton / permutation / interpolation-matrix outputs
(matlab/RBRP.m:2, matlab/driver.m:83-94).

PyTorch is used only to hold device memory and streams; every computation
runs in ``librbrp.so``.
"""

from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import numpy as np

from . import _capi


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("rbrp: no GPU device is visible to HIP")
    return torch


def colmajor(A, keep_fp32: bool = False):
    """A column-major (stride (1, ld)) FP64 device view of ``A`` (copies only if needed).

    With ``keep_fp32`` an FP32 device matrix stays FP32."""
    torch = _torch()
    if keep_fp32 and A.is_cuda and A.dtype == torch.float32:
        pass
    elif not (A.is_cuda and A.dtype == torch.float64):
        A = A.to(device="cuda", dtype=torch.float64)
    m, n = A.shape
    if A.stride(0) == 1 and (n <= 1 or A.stride(1) >= max(m, 1)):
        return A
    return A.t().contiguous().t()


def leading_dim(A) -> int:
    m, n = A.shape
    return int(A.stride(1)) if n > 1 else max(m, 1)


class Handle:
    """Owns a library handle bound to a HIP stream (default: torch's current stream)."""

    def __init__(self, stream=None):
        torch = _torch()
        self._lib = _capi.lib()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p()
        _capi.check("rbrp_create", self._lib.rbrp_create(ctypes.byref(h),
                                                        ctypes.c_void_p(self.stream.cuda_stream)))
        self._h = h

    @property
    def ptr(self):
        return self._h

    def set_timing(self, enable):
        """Per-phase event timing: False/0 off, True/1 every phase, 2 only the
        projection and update phases (the least added to the stream)."""
        level = int(enable)
        _capi.check("rbrp_set_timing", self._lib.rbrp_set_timing(self._h, level))

    def reset_stats(self):
        _capi.check("rbrp_reset_phase_stats", self._lib.rbrp_reset_phase_stats(self._h))

    def stats(self) -> dict:
        s = _capi.RbrpPhaseStats()
        _capi.check("rbrp_get_phase_stats", self._lib.rbrp_get_phase_stats(self._h, ctypes.byref(s)))
        return {ph: dict(ms=s.ms[q], launches=int(s.launches[q]), flops=s.flops[q], bytes=s.bytes[q])
                for q, ph in enumerate(_capi.PHASES)}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.rbrp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


@dataclasses.dataclass
class ColID:
    k: int
    perm: "object"          # torch int32 (n,), device
    X: "object"             # torch float64 (k, n-k) column-major view, device (None if skipped)
    ranks: np.ndarray       # cumulative rank per block
    error: np.ndarray       # squared relative residual per block
    t_ranks: np.ndarray     # cumulative seconds per block
    t_total: float
    err_final: float        # non-squared final relative residual
    draws: int              # uniforms consumed

    @property
    def skeleton(self):
        return self.perm[: self.k]

    @property
    def redundant(self):
        return self.perm[self.k:]


def rbrp_colid(A, b: int = 64, tol: float = 0.0, seed: int = 0, offset: int = 0, max_rank: int = 0,
               filter: str = "sampled", precision: str = "fp64", want_x: bool = True,
               handle: Optional[Handle] = None, noalloc: bool = False) -> ColID:
    """RBRP column ID of the device matrix ``A``; see the module docstring.

    ``A`` is m x n FP64, or FP32 (kept in FP32 and passed to ``rbrp_colid_s``).
    ``precision`` is "fp64" (the reference's arithmetic) or "fp32" (FP32-stored
    residual with FP32 MFMA projection / update; FP64 sampled-block
    factorisation, re-orthogonalisation, stopping sum and solve).  ``noalloc``
    calls ``rbrp_colid_d_noalloc`` (FP64 input; no device allocation inside the
    call, workspace from the library's static scratch)."""
    torch = _torch()
    A = colmajor(A, keep_fp32=not noalloc)
    fn = "rbrp_colid_s" if A.dtype == torch.float32 else "rbrp_colid_d"
    if noalloc:
        fn = "rbrp_colid_d_noalloc"
    m, n = A.shape
    own = handle is None
    h = Handle() if own else handle
    try:
        p = min(m, n)
        kmax = min(max_rank, p) if max_rank else p
        ldx = max(kmax, 1)
        perm = torch.empty(n, dtype=torch.int32, device=A.device)
        xbuf = torch.empty((max(n, 1), ldx), dtype=torch.float64, device=A.device) if want_x else None
        cap = max(kmax, 1) + 1
        ranks = np.zeros(cap, dtype=np.int64)
        error = np.zeros(cap, dtype=np.float64)
        t_ranks = np.zeros(cap, dtype=np.float64)
        opts = _capi.RbrpOpts(b=int(b), tol=float(tol), max_rank=int(max_rank), seed=int(seed),
                              stream_offset=int(offset), precision=_capi.PRECISIONS[precision],
                              filter=_capi.FILTERS[filter])
        res = _capi.RbrpResult()
        res.perm = perm.data_ptr()
        res.X = xbuf.data_ptr() if want_x else None
        res.ldx = ldx
        res.trace_cap = cap
        res.ranks = ranks.ctypes.data
        res.error = error.ctypes.data
        res.t_ranks = t_ranks.ctypes.data
        code = getattr(h._lib, fn)(h.ptr, ctypes.c_void_p(A.data_ptr()), m, n, leading_dim(A),
                                   ctypes.byref(opts), ctypes.byref(res))
        _capi.check(fn, code)
        k = int(res.k)
        X = xbuf[: n - k, :k].t() if want_x else None
        nb = int(res.nblocks)
        return ColID(k=k, perm=perm, X=X, ranks=ranks[:nb].copy(), error=error[:nb].copy(),
                     t_ranks=t_ranks[:nb].copy(), t_total=float(res.t_total),
                     err_final=float(res.err_final), draws=int(res.draws))
    finally:
        if own:
            h.close()


def RBRP(m: int, n: int, A, blk: int, seed: int = 0, offset: int = 0):
    """Mirror of ``RBRP(m, n, A, blk, ranks, error, t_ranks, t_rbrp, flops)`` (gpu/rbrp.hpp:6-8).

    Returns ``(ranks, error, t_ranks, t_rbrp, flops)``; like the reference it
    runs to ``min(m, n)`` with no tolerance and leaves ``A`` untouched.
    ``flops`` is the algorithmic FP64 count of the trailing updates,
    ``sum_t 4 m b_t (n - i_t - b_t)`` (the reference declares but never sets it).
    """
    torch = _torch()
    A = torch.as_strided(A, (m, n), (1, m)) if A.dim() == 1 else A
    r = rbrp_colid(A[:m, :n], b=blk, tol=0.0, seed=seed, offset=offset, want_x=False)
    prev = np.concatenate([[0], r.ranks[:-1]]) if len(r.ranks) else np.zeros(0)
    bts = r.ranks - prev
    flops = float(np.sum(4.0 * m * bts * (n - r.ranks)))
    return (list(map(int, r.ranks)), list(map(float, r.error)), list(map(float, r.t_ranks)),
            r.t_total, flops)


def _batch_setup(As, b, tol, seeds, offsets, max_rank, filter, want_x):
    """Descriptors and output buffers of a batch (keeps the tensors alive in the return)."""
    torch = _torch()
    nprob = len(As)
    seeds = list(range(nprob)) if seeds is None else list(seeds)
    offsets = [0] * nprob if offsets is None else list(offsets)
    probs = (_capi.RbrpProblem * max(nprob, 1))()
    keep, perms, xbufs = [], [], []
    for q, A in enumerate(As):
        A = colmajor(A)
        m, n = A.shape
        pmn = min(m, n)
        kmax = min(max_rank, pmn) if max_rank else pmn
        ldx = max(kmax, 1)
        perm = torch.empty(max(n, 1), dtype=torch.int32, device=A.device)
        xb = torch.empty((max(n, 1), ldx), dtype=torch.float64, device=A.device) if want_x else None
        keep.append(A)
        perms.append(perm)
        xbufs.append(xb)
        probs[q].A = A.data_ptr()
        probs[q].m, probs[q].n, probs[q].lda = m, n, leading_dim(A)
        probs[q].seed, probs[q].stream_offset = int(seeds[q]), int(offsets[q])
        probs[q].perm = perm.data_ptr()
        probs[q].X = xb.data_ptr() if want_x else None
        probs[q].ldx = ldx
    opts = _capi.RbrpOpts(b=int(b), tol=float(tol), max_rank=int(max_rank), seed=0, stream_offset=0,
                          precision=0, filter=_capi.FILTERS[filter])
    return probs, opts, keep, perms, xbufs


class Batch:
    """A batch of device FP64 problems with its descriptors and output buffers,
    built once and run any number of times (``run``).  Problem q uses the
    Philox stream (seeds[q], offsets[q]) (default (q, 0)); b, tol, max_rank and
    filter are shared."""

    def __init__(self, As, b: int = 16, tol: float = 0.0, seeds=None, offsets=None,
                 max_rank: int = 0, filter: str = "sampled", want_x: bool = True):
        self.nprob = len(As)
        self.want_x = want_x
        (self.probs, self.opts, self._keep, self.perms,
         self.xbufs) = _batch_setup(As, b, tol, seeds, offsets, max_rank, filter, want_x)
        self.res = (_capi.RbrpProblemResult * max(self.nprob, 1))()

    def run(self, mode: str = "persistent", handle: Optional[Handle] = None) -> np.ndarray:
        """One batched call; returns the per-problem status array."""
        own = handle is None
        h = Handle() if own else handle
        try:
            _capi.check("rbrp_colid_batched_d", h._lib.rbrp_colid_batched_d(
                h.ptr, self.nprob, self.probs, ctypes.byref(self.opts), self.res,
                _capi.BATCH_MODES[mode]))
            h.stream.synchronize()
        finally:
            if own:
                h.close()
        return np.array([int(self.res[q].status) for q in range(self.nprob)], dtype=np.int32)

    def results(self):
        """Per-problem ``ColID`` of the last run (per-block traces empty)."""
        out = []
        for q in range(self.nprob):
            k, n = int(self.res[q].k), int(self._keep[q].shape[1])
            X = self.xbufs[q][: n - k, :k].t() if self.want_x else None
            out.append(ColID(k=k, perm=self.perms[q][:n], X=X, ranks=np.zeros(0, np.int64),
                             error=np.zeros(0), t_ranks=np.zeros(0), t_total=0.0,
                             err_final=float(self.res[q].err_final),
                             draws=int(self.res[q].draws)))
        return out


def rbrp_colid_batched(As, b: int = 16, tol: float = 0.0, seeds=None, offsets=None,
                       max_rank: int = 0, filter: str = "sampled", mode: str = "auto",
                       want_x: bool = True, handle: Optional[Handle] = None):
    """Batched RBRP column IDs of the device FP64 matrices ``As`` (a list).

    ``mode``: "persistent" (one persistent launch with a device work queue,
    replayed from a hipGraph), "loop" (one ``rbrp_colid_d`` per problem) or
    "auto".  Returns the list of per-problem ``ColID`` and the status array.
    """
    batch = Batch(As, b, tol, seeds, offsets, max_rank, filter, want_x)
    status = batch.run(mode, handle)
    return batch.results(), status


def batched_tune(As, b: int = 16, tol: float = 0.0, seeds=None, offsets=None, max_rank: int = 0,
                 filter: str = "sampled", handle: Optional[Handle] = None) -> int:
    """Run the persistent batched launch once per candidate launch shape (update
    row tiles x workgroups per CU) and keep the fastest on ``handle``; returns
    its workgroups per CU (``batched_launch_shape`` gives both)."""
    probs, opts, keep, perms, xbufs = _batch_setup(As, b, tol, seeds, offsets, max_rank, filter,
                                                   True)
    pick = ctypes.c_int(0)
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_batched_tune", h._lib.rbrp_batched_tune(
            h.ptr, len(As), probs, ctypes.byref(opts), ctypes.byref(pick)))
    finally:
        if own:
            h.close()
    del keep, perms, xbufs
    return int(pick.value)


def batched_launch_shape(As, b: int = 16, tol: float = 0.0, max_rank: int = 0,
                         filter: str = "sampled", handle: Optional[Handle] = None):
    """``(workgroups per CU, update row tiles)`` that persistent calls for this
    batch shape use on ``handle`` (the tuner's pick, or the default)."""
    probs, opts, keep, perms, xbufs = _batch_setup(As, b, tol, None, None, max_rank, filter, True)
    wg, tile = ctypes.c_int(0), ctypes.c_int(0)
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_batched_launch_shape", h._lib.rbrp_batched_launch_shape(
            h.ptr, len(As), probs, ctypes.byref(opts), ctypes.byref(wg), ctypes.byref(tile)))
    finally:
        if own:
            h.close()
    del keep, perms, xbufs
    return int(wg.value), int(tile.value)


def uniforms(seed: int, offset: int, draw0: int, count: int, handle: Optional[Handle] = None):
    """Uniforms ``draw0 .. draw0+count-1`` of the Philox stream, computed on the device."""
    torch = _torch()
    out = torch.empty(count, dtype=torch.float64, device="cuda")
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_uniforms_d", h._lib.rbrp_uniforms_d(h.ptr, seed, offset, draw0, count,
                                                              ctypes.c_void_p(out.data_ptr())))
        h.stream.synchronize()
    finally:
        if own:
            h.close()
    return out


def gaussian_sketch(rows: int, cols: int, scale: float, seed: int = 0, offset: int = 0,
                    handle: Optional[Handle] = None):
    """Device Gaussian sketch (rows x cols, column-major) from the Philox stream (seed, offset)."""
    torch = _torch()
    buf = torch.empty((cols, rows), dtype=torch.float64, device="cuda")
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_gaussian_sketch_d", h._lib.rbrp_gaussian_sketch_d(
            h.ptr, seed, offset, rows, cols, float(scale), ctypes.c_void_p(buf.data_ptr()), rows))
        h.stream.synchronize()
    finally:
        if own:
            h.close()
    return buf.t()


def _sketch_args(A, G):
    A = colmajor(A)
    if G is not None:
        G = colmajor(G)
        return A, ctypes.c_void_p(G.data_ptr()), leading_dim(G)
    return A, None, 0


def _sketch_id(cname: str, host_idx: bool, A, rank: int, seed: int, offset: int, G, handle):
    """One sketch ID call through the C ABI: ``(sk, rd, T)`` with ``sk``/``rd`` host
    int arrays (``host_idx``, the CPQR forms) or device tensors (the LUPP forms)."""
    torch = _torch()
    A, gp, ldg = _sketch_args(A, G)
    m, n = A.shape
    if host_idx:
        sk = np.zeros(rank, dtype=np.int32)
        rd = np.zeros(max(n - rank, 1), dtype=np.int32)
        skp, rdp = ctypes.c_void_p(sk.ctypes.data), ctypes.c_void_p(rd.ctypes.data)
    else:
        sk = torch.empty(rank, dtype=torch.int32, device=A.device)
        rd = torch.empty(max(n - rank, 1), dtype=torch.int32, device=A.device)
        skp, rdp = ctypes.c_void_p(sk.data_ptr()), ctypes.c_void_p(rd.data_ptr())
    Tb = torch.empty((max(n - rank, 1), rank), dtype=torch.float64, device=A.device)
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check(cname, getattr(h._lib, cname)(
            h.ptr, ctypes.c_void_p(A.data_ptr()), m, n, leading_dim(A), rank, seed, offset, gp, ldg,
            skp, rdp, ctypes.c_void_p(Tb.data_ptr()), rank))
    finally:
        if own:
            h.close()
    return sk, rd[: n - rank], Tb[: n - rank].t()


def RandCPQR_OS(A, rank: int, seed: int = 0, offset: int = 0, G=None,
                handle: Optional[Handle] = None):
    """Mirror of ``RandCPQR_OS(A, m, n, rank, h_sk, h_rd, T, flops)`` (gpu/rid.hpp:19-20).

    Returns ``(sk, rd, T)``: host int arrays (as the reference's ``h_sk``/``h_rd``)
    and the device ``T`` (rank x (n - rank)).  ``G`` optionally supplies the
    2*rank x m sketch; otherwise it is drawn from the Philox stream (seed, offset).
    """
    return _sketch_id("rbrp_randcpqr_os_d", True, A, rank, seed, offset, G, handle)


def RandCPQR(A, rank: int, seed: int = 0, offset: int = 0, G=None,
             handle: Optional[Handle] = None):
    """Mirror of ``RandCPQR(A, m, n, rank, h_sk, h_rd, T, flops)`` (gpu/rid.hpp:16-17).

    As :func:`RandCPQR_OS` without oversampling: ``G`` is rank x m and ``T`` comes
    from the sketch's own triangular factor (gpu/randCPQR.cu:15-67).
    """
    return _sketch_id("rbrp_randcpqr_d", True, A, rank, seed, offset, G, handle)


def RandLUPP_OS(A, rank: int, seed: int = 0, offset: int = 0, G=None,
                handle: Optional[Handle] = None):
    """Mirror of ``RandLUPP_OS(A, m, n, rank, sk, rd, T, flops)`` (gpu/rid.hpp:12-13).

    Returns ``(sk, rd, T)`` as device tensors (the reference's ``sk``/``rd`` are
    device pointers too).
    """
    return _sketch_id("rbrp_randlupp_os_d", False, A, rank, seed, offset, G, handle)


def RandLUPP(A, rank: int, seed: int = 0, offset: int = 0, G=None,
             handle: Optional[Handle] = None):
    """Mirror of ``RandLUPP(A, m, n, rank, sk, rd, T, flops)`` (gpu/rid.hpp:9-10).

    As :func:`RandLUPP_OS` without oversampling: ``G`` is rank x m and
    ``T = (L21 L11^{-1})^T`` from the LU factor (gpu/randLUPP.cu:13-107).
    """
    return _sketch_id("rbrp_randlupp_d", False, A, rank, seed, offset, G, handle)


def CPQR(A, ranks, handle: Optional[Handle] = None):
    """Mirror of ``CPQR(m, n, A, ranks, err, t_ranks, t_cpqr, flops)`` (gpu/id_cpqr.hpp:6-8).

    Full-matrix column-pivoted QR of ``A`` (device, FP64, m x n), then for each
    rank r in ``ranks`` the interpolation matrix R11^{-1} R12 and its squared
    relative error (gpu/id_cpqr.cu:30-172).  Returns ``(error, t_ranks, t_cpqr,
    flops, perm)``: host lists of len(ranks), the factorization's seconds, the
    flop count and the device pivot order (int32, n entries).
    """
    torch = _torch()
    A = colmajor(A)
    m, n = A.shape
    rk = np.ascontiguousarray(np.asarray(ranks, dtype=np.int32))
    err = np.zeros(max(len(rk), 1))
    tr = np.zeros(max(len(rk), 1))
    tq, fl = ctypes.c_double(), ctypes.c_double()
    perm = torch.empty(n, dtype=torch.int32, device=A.device)
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_cpqr_id_d", h._lib.rbrp_cpqr_id_d(
            h.ptr, ctypes.c_void_p(A.data_ptr()), m, n, leading_dim(A), len(rk),
            ctypes.c_void_p(rk.ctypes.data), ctypes.c_void_p(err.ctypes.data),
            ctypes.c_void_p(tr.ctypes.data), ctypes.byref(tq), ctypes.byref(fl),
            ctypes.c_void_p(perm.data_ptr())))
    finally:
        if own:
            h.close()
    return list(err[: len(rk)]), list(tr[: len(rk)]), tq.value, fl.value, perm


def RandAdapLUPP(A, tol: float = 1e-8, blk: int = 16, seed: int = 0, offset: int = 0,
                 handle: Optional[Handle] = None):
    """Mirror of ``RandAdapLUPP(A, m, n, sk, rd, T, rank, flops, tol, blk)``
    (gpu/rid.hpp:5-7; the semantics of blocked_rlupp_id, python/cache.ipynb
    c8:80-143, in column form): adaptive-rank randomised LUPP ID of the device
    matrix ``A`` (FP64, m x n), one Gaussian sketch of ``blk`` rows per block,
    stopping when the sketched error estimate is at most ``tol * ||A||_F``.
    Returns ``(sk, rd, T, rank, err_est, flops)``: device int32 skeleton and
    redundant indices, ``T`` (rank x (n - rank)) with ``A[:, rd] ~ A[:, sk] @ T``,
    the last relative estimate and the arithmetic count."""
    torch = _torch()
    A = colmajor(A)
    m, n = A.shape
    p = min(m, n)
    sk = torch.empty(max(n, 1), dtype=torch.int32, device=A.device)
    rd = torch.empty(max(n, 1), dtype=torch.int32, device=A.device)
    Tb = torch.empty((max(n, 1), max(p, 1)), dtype=torch.float64, device=A.device)
    rank, est, fl = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_randadaplupp_d", h._lib.rbrp_randadaplupp_d(
            h.ptr, ctypes.c_void_p(A.data_ptr()), m, n, leading_dim(A), float(tol), int(blk),
            int(seed), int(offset), ctypes.c_void_p(sk.data_ptr()), ctypes.c_void_p(rd.data_ptr()),
            ctypes.c_void_p(Tb.data_ptr()), max(p, 1), ctypes.byref(rank), ctypes.byref(est),
            ctypes.byref(fl)))
    finally:
        if own:
            h.close()
    k = int(rank.value)
    T = Tb[: n - k, :k].t()
    return sk[:k], rd[: n - k], T, k, float(est.value), float(fl.value)


def interp_os(Y, P, k: int, handle: Optional[Handle] = None):
    """Mirror of ``Compute_interpolation_matrix`` (gpu/interp.cu:45-105): T (k x (n-k))."""
    torch = _torch()
    Y = colmajor(Y)
    l, n = Y.shape
    P = torch.as_tensor(P, dtype=torch.int32, device=Y.device).contiguous()
    Tb = torch.empty((max(n - k, 1), k), dtype=torch.float64, device=Y.device)
    own = handle is None
    h = Handle() if own else handle
    try:
        _capi.check("rbrp_interp_os_d", h._lib.rbrp_interp_os_d(
            h.ptr, ctypes.c_void_p(Y.data_ptr()), l, n, leading_dim(Y), k,
            ctypes.c_void_p(P.data_ptr()), ctypes.c_void_p(Tb.data_ptr()), k))
    finally:
        if own:
            h.close()
    return Tb[: n - k].t()
